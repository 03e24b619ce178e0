"""Builds weibull_golden.cc with the host's g++ / libstdc++ and writes its
output to weibull_default_engine.txt (the golden vector of the Weibull
threshold draws; see the .cc file)."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "weibull_golden")
        subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "weibull_golden.cc"),
                        "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(HERE, "weibull_default_engine.txt"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
