"""fock_amd -- MI355X restricted Hartree-Fock engine (host side over libfock.so).

Modules:
    molecule  basis sets, xyz input, geometry generators (the reference's inputs)
    abi       ctypes binding of include/fock.h
    scf       mirror of the reference's `scf` crate API (try_scf, HartreeFockResult, ...)
"""
from . import molecule  # noqa: F401

__all__ = ["molecule", "abi", "scf"]
