"""Algorithmic FLOP model of the chemistry step, for roofline accounting.

Convention: add, sub, mul, div = 1 flop; fma = 2; compares, selects and
data movement = 0; exp and log count the operations of the project's own FP64
implementations (tiger_nga_amd/csrc/fp64_math.h): exp 33, log 32.  Per cell
step:

  F = (nfe + nfe_dq) * F_rhs + nlu * F_lu + nni * F_solve + nfe_dq * 3 neq + nst * F_step

* F_rhs   one right-hand side (finitechem_compute_rhs, reference
          combustion/finitechem.f90:93-175, with PRODRATES and COMPTHERMODATA),
          counted from the mechanism tables as the kernel evaluates them;
* F_lu    forming M = I - gamma J (2 neq^2) plus LU with partial pivoting
          (2/3 neq^3);
* F_solve forward and back substitution (2 neq^2) plus the Newton update
          (4 neq);
* 3 neq   per difference-quotient column (scaled difference of two RHS);
* F_step  Nordsieck predict / correct / error norms, taken as 30 neq.
"""
from __future__ import annotations

EXP = 33.0
LOG = 32.0
ARRH = 1 + 2 + EXP + 1            # e*irt, fma(b, lt, .), exp, A*
TROE = 2 + (LOG + 2) + 3 + (2 + 1 + EXP) + 4  # one direction of GETLINDRATECOEFF


def rhs_flops(info: dict) -> float:
    nsp = info["nsp"]
    f = 0.0
    f += nsp * (1 + 20 + 1 + 3 + 1)   # rho sum, thermo h+cp, C = y/W, cp_mix terms, ctot
    f += 1                            # T = rhoT / rho
    f += LOG + 1 + 1                  # ln T, R T, 1/(R T)
    f += 2 * info["n_eff_corr"]       # third-body corrections
    f += ARRH * (info["nrxn"] + info["n_reversible"])
    f += info["n_conc_factors"] + info["n_third_body_rxn"] + info["n_reversible"]  # W products, q = Wf - Wb
    f += info["n_fc_terms"] * (1 + EXP + 2)             # Fc terms
    f += info["n_falloff"] * (LOG + 1 + 2 + 2)           # ln Fc, log10, n, c
    f += (ARRH + TROE) * (info["n_falloff"] + info["n_rev_falloff"])
    f += 2 * info["n_nu"]             # CDOT gather
    f += nsp * (1 + 2) + 1            # rhs_k = CDOT W, -sum h rhs, / cp
    return f


def rhs_flops_cached(info: dict) -> float:
    """One right-hand side at a temperature whose factors are already known
    (the species columns of the guided solve's split Jacobian): no ln T,
    thermo polynomials, Arrhenius terms or Troe centres; the fall-off
    blending (Pr, F) is still evaluated."""
    nsp = info["nsp"]
    return (rhs_flops(info) - nsp * 20 - LOG - ARRH * (info["nrxn"] + info["n_reversible"])
            - info["n_fc_terms"] * (1 + EXP + 2) - info["n_falloff"] * LOG
            - ARRH * (info["n_falloff"] + info["n_rev_falloff"]))


def step_flops(info: dict, stats: dict) -> float:
    """Total algorithmic flops of one call from its integrator counters."""
    n = info["nsp"] + 1
    f_rhs = rhs_flops(info)
    f_lu = 2.0 * n * n + 2.0 / 3.0 * n ** 3
    f_solve = 2.0 * n * n + 4.0 * n
    return ((stats["nfe"] + stats["nfe_dq"]) * f_rhs + stats["nlu"] * f_lu + stats["nni"] * f_solve
            + stats["nfe_dq"] * 3.0 * n + stats["nst"] * 30.0 * n)


def step_bytes(info: dict, cells: int) -> float:
    """Algorithmic HBM bytes of one call: read rho, Y, T and write Y, T, HR and
    three source terms per cell, 8 bytes each."""
    nsp = info["nsp"]
    return 8.0 * cells * ((nsp + 2) + (nsp + 1) + 4)


def diffusion_bytes(nsp: int, cells: int) -> float:
    """Algorithmic HBM bytes of one finitechem_diffusion call on the interior
    cells, 8 bytes each: read rho, Y_k and T, and the diffusivities of the
    species and T; write diff_corr (species and T) and the three diff_gp
    planes.  Stencil neighbours and intermediate planes are not counted."""
    return 8.0 * cells * (1 + 3 * (nsp + 1) + 3)


def properties_flops(nsp: int) -> float:
    """Algorithmic flops of one cell of the properties kernel (same convention;
    sqrt = 1): W_mix (3 nsp), species viscosities (omegamu 16 + 3 per species),
    Wilke double sum (10 per pair: ratio, sqrt, two products, square, the
    weighted term and its accumulation), conductivity and the two mixture sums
    (12 per species), density and outputs (5 + nsp)."""
    return 3 * nsp + 19 * nsp + 10 * nsp * nsp + 12 * nsp + 5 + nsp


def guided_flops(info: dict, sstats: dict, nsub: int = 0, split: bool = False) -> float:
    """Algorithmic flops of the surrogate step's guided solve (before the plain
    fallback) from its counters: every right-hand side (Newton and
    difference-quotient columns), each factorisation of I - gamma h J counted
    as an LU (2 n^2 + 2/3 n^3: the kernel forms the Gauss-Jordan inverse, n^3
    multiply-adds, and the excess over the LU is not counted), one filtered
    error estimate per sub-step taken, a solve (2 n^2, as two triangular
    sweeps or one product with the inverse) and update per Newton iteration,
    and the enthalpy check / element projection (taken as 60 neq per cell).  Without the factorisation counter, one LU and one
    estimate per counted sub-step (or nsub per attempted cell).  split: each
    Jacobian's nsp species columns are fixed-temperature evaluations
    (rhs_flops_cached)."""
    n = info["nsp"] + 1
    cells = sstats["accepted"] + sstats["rejected"]
    f_lu = 2.0 * n * n + 2.0 / 3.0 * n ** 3
    f_solve = 2.0 * n * n + 4.0 * n
    subs = sstats.get("substeps") or cells * nsub
    lus = sstats.get("factorizations") or subs
    jacs = sstats.get("jacobians") or cells
    f = (sstats["rhs_evals"] * rhs_flops(info) + lus * f_lu + subs * f_solve
         + sstats["newton_iters"] * f_solve + cells * 60.0 * n)
    if split:
        f -= jacs * info["nsp"] * (rhs_flops(info) - rhs_flops_cached(info))
    return f
