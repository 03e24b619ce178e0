"""akantu_amd: an MI355X-native explicit-dynamics engine for the central-difference
path of Akantu's SolidMechanicsModel.

The compute path is the HIP library ``akantu_amd/libakx.so`` (C ABI declared in
``include/akx.h``), reached from Python through :mod:`akantu_amd.akx` and from
C++ through the ``akantu::`` façade in ``akantu_amd/facade``.  Importing the
package itself loads nothing; :mod:`akantu_amd.akx` raises if the library is
missing.
"""
__all__ = ["mesh", "akx"]
