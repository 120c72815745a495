"""Writes the golden vectors of the mechanism routines.

Runs the reference mechanism files themselves -- combustion/HYDROGENF.f90,
grimech30F.f90 and DME39F.f90, compiled unchanged with flang by
``make -C oracle ref`` into oracle/_ref/ -- on seeded states and stores their
outputs as fixtures (``<mech>_mech.npz``):

* PRODRATES(CDOT, W, K, C, M, TEMP, PRESSURE)   (HYDROGENF.f90:4-277)
  -> per-slot rate coefficients K, rates W, third-body sums M, CDOT
* COMPTHERMODATA(H, CP, T)                       (HYDROGENF.f90:514-673)
  on a temperature ladder (H, CP) and at every state's T (H_state, CP_state:
  with CDOT they fix finitechem_compute_rhs's temperature row)
* GETMOLARMASS / GETSPECIESNAMES / GETNSPECIES / GETNREACTIONS
* GETMUCOEFF / GETKOVEREPS                       (HYDROGENF.f90:453-496)

States cover T in [300, 3000] K across the 1000 K polynomial switch (and
exactly at it), 1 and 10 atm, log-uniform concentrations, exact zeros (M = 0
takes GETLINDRATECOEFF's P/(R T) branch) and small negative concentrations
(allowed inside the solver).  Run in the build container only:
``python tests/golden/make_golden.py``.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_%s.so")
MECHS = ("h2", "gri30", "dme39", "ethyred")
NSTATE = 24
NTHERMO = 41

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a):
    return a.ctypes.data_as(dp)


def load(mech):
    lib = ctypes.CDLL(REF_LIB % mech, mode=os.RTLD_LOCAL)
    n = ctypes.c_int()
    lib.getnspecies_(ctypes.byref(n))
    nsp = n.value
    lib.getnreactions_(ctypes.byref(n))
    nslot = n.value
    return lib, nsp, nslot


def states(nsp, seed):
    rng = np.random.default_rng(seed)
    T = rng.uniform(300.0, 3000.0, NSTATE)
    T[:4] = [1000.0, np.nextafter(1000.0, 2000.0), 300.0, 3000.0]
    P = np.where(rng.random(NSTATE) < 0.5, 101325.0, 1013250.0)
    C = 10.0 ** rng.uniform(-12.0, -1.5, (NSTATE, nsp))
    C[4, :] = 0.0                              # M = 0 -> P/(RT) fall-off branch
    C[5, rng.random(nsp) < 0.5] = 0.0          # sparse state
    C[6, :] = -1e-9 * rng.random(nsp)          # negative (solver iterates)
    C[7, rng.random(nsp) < 0.3] *= -1e-3
    return T, P, C


def main(argv=None):
    only = set(argv or sys.argv[1:])
    for i, mech in enumerate(MECHS):
        if only and mech not in only:
            continue
        lib, nsp, nslot = load(mech)
        # third-body sums: the mechanism declares M(n) locally; probe size from the header
        nm = 64
        W_mol = np.zeros(nsp)
        lib.getmolarmass_(_ptr(W_mol))
        mucoeff, koveps = np.zeros(nsp), np.zeros(nsp)
        lib.getmucoeff_(_ptr(mucoeff))
        lib.getkovereps_(_ptr(koveps))
        names = ctypes.create_string_buffer(20 * nsp)
        lib.getspeciesnames_(names, ctypes.c_long(20))
        names = [names.raw[20 * k:20 * k + 20].decode().strip() for k in range(nsp)]
        T, P, C = states(nsp, 1234 + i)
        CDOT = np.zeros((NSTATE, nsp))
        K = np.zeros((NSTATE, nslot))
        W = np.zeros((NSTATE, nslot))
        M = np.zeros((NSTATE, nm))
        for s in range(NSTATE):
            cd, w, k, m = np.zeros(nsp), np.zeros(nslot), np.zeros(nslot), np.zeros(nm)
            c = C[s].copy()
            t, p = ctypes.c_double(T[s]), ctypes.c_double(P[s])
            lib.prodrates_(_ptr(cd), _ptr(w), _ptr(k), _ptr(c), _ptr(m), ctypes.byref(t), ctypes.byref(p))
            CDOT[s], K[s], W[s], M[s] = cd, k, w, m
        Tt = np.concatenate([np.linspace(250.0, 3500.0, NTHERMO - 3),
                             [1000.0, np.nextafter(1000.0, 0.0), np.nextafter(1000.0, 2000.0)]])
        H = np.zeros((len(Tt), nsp))
        CP = np.zeros((len(Tt), nsp))
        for s, tv in enumerate(Tt):
            h, cp = np.zeros(nsp), np.zeros(nsp)
            lib.compthermodata_(_ptr(h), _ptr(cp), ctypes.byref(ctypes.c_double(tv)))
            H[s], CP[s] = h, cp
        Hs = np.zeros((NSTATE, nsp))
        CPs = np.zeros((NSTATE, nsp))
        for s in range(NSTATE):
            h, cp = np.zeros(nsp), np.zeros(nsp)
            lib.compthermodata_(_ptr(h), _ptr(cp), ctypes.byref(ctypes.c_double(T[s])))
            Hs[s], CPs[s] = h, cp
        out = os.path.join(HERE, "%s_mech.npz" % mech)
        np.savez_compressed(out, C=C, T=T, P=P, CDOT=CDOT, K=K, W=W, M=M, T_thermo=Tt, H=H, CP=CP,
                            H_state=Hs, CP_state=CPs, W_mol=W_mol, mucoeff=mucoeff, koveps=koveps,
                            names=np.array(names), nslot=nslot)
        print("%s: %d species, %d slots -> %s (%d bytes)" % (mech, nsp, nslot, out, os.path.getsize(out)))


if __name__ == "__main__":
    sys.exit(main())
