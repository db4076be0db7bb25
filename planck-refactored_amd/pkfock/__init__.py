"""pkfock — MI355X-native direct-SCF Fock-build engine (HIP, gfx950).

engine.FockEngine  : the C ABI of include/pk_fock.h (J/K, G, Schwarz, ERIs)
host.Molecule      : basis, one-electron integrals, device RHF SCF (include/pk_scf.h)
molecules          : synthetic geometries of the benchmark workloads
"""
