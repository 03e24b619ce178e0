// Writes the golden vector of the reference's Weibull thresholds: the
// std::default_random_engine of RandomGenerator (aka_random_generator.hh:72-96)
// seeded with the run's seed, values base + std::weibull_distribution<double>
// (shape, scale) in order (RandomParameter::setValues, :230-234), for the
// notched-plate workload's law: Weibull(m = 5, scale = 1 MPa), base 0.
// Built and run by make_weibull_golden.py.
#include <cstdio>
#include <random>

int main() {
  const long seeds[] = {5489, 1, 42};
  for (long s : seeds) {
    std::default_random_engine engine(static_cast<unsigned>(s));
    std::weibull_distribution<double> w(5.0, 1e6);
    std::printf("seed %ld\n", s);
    for (int k = 0; k < 16; ++k) std::printf("%.17g\n", 0.0 + w(engine));
  }
  return 0;
}
