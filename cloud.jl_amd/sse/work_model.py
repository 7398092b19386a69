"""Algorithmic work of the residual kernels (for roofline reporting).

Counts are per element and per residual, from the operator sparsity of the
reference element, for the minimal algorithm: each volume node pair (i < j) of
the union sparsity of S_1..S_d evaluated once, each nonzero of C once, each
facet node once, sum-factorised V/V'. Divides, square roots, logarithms,
exponentials and powers are counted separately and not as flops; a fused
multiply-add counts as 2 flops.

Bytes are compulsory HBM traffic: each input read once, each output written
once; reference-element operators are cache resident and not counted.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .reference_approximation import ReferenceApproximation, flux_differencing_operators

# flops of one entropy-conservative two-point flux evaluation (Ranocha), by
# dimension: velocities/pressures, two logarithmic means and the d x (d+2) tensor
EC_FLUX_FLOPS = {1: 50, 2: 75, 3: 100}


def sum_factorized_v_flops(ref: ReferenceApproximation) -> int:
    """Flops of one V (or V') application to one variable."""
    wp = ref.warped
    if wp is None:
        return 2 * ref.N_q * ref.N_p if ref.V is not None else 0
    p = wp.p
    if wp.C is None:
        M1, M2 = wp.M
        st1 = sum(2 * M2 * (p + 1 - i) for i in range(p + 1))
        st2 = 2 * M1 * M2 * (p + 1)
        return st1 + st2
    M1, M2, M3 = wp.M
    st1 = sum(2 * M3 * (p + 1 - i - j) for i in range(p + 1) for j in range(p + 1 - i))
    st2 = sum(2 * M2 * M3 * (p + 1 - i) for i in range(p + 1))
    st3 = 2 * M1 * M2 * M3 * (p + 1)
    return st1 + st2 + st3


@dataclass
class KernelWork:
    flops: float
    bytes: float
    logs: float
    divides: float


def residual_work(ref: ReferenceApproximation, N_c: int) -> dict:
    """Per-element work of the two kernels of the flux-differencing residual."""
    d = ref.d
    S, S_sparse, C, C_kind = flux_differencing_operators(ref)
    union = np.triu(np.abs(S).sum(0), 1) != 0
    pairs = int(union.sum())
    nnzC = 0 if C is None else int((C != 0).sum())
    nnzR = int((ref.R != 0).sum())
    N_p, N_q, N_f = ref.N_p, ref.N_q, ref.N_f
    cF = EC_FLUX_FLOPS[d]
    v = sum_factorized_v_flops(ref)
    # phase 1: V u, c2e, WJ scale, V', WADG (V, scale, V'), V, R, e2c at volume and facet nodes
    nv = 5 if ref.V is not None else 0
    c2e = 10 + 4 * d
    e2c = 12 + 4 * d
    f1 = (nv * v * N_c + N_q * (c2e + 2 * N_c) + N_q * N_c + 2 * nnzR * N_c
          + (N_q + N_f) * e2c)
    b1 = 8 * (N_p * N_c + N_q + N_q * N_c + N_f * N_c)
    # phase 2: interface flux, Hadamard sum, facet correction, lift, V' + WADG
    per_pair = cF + 3 * d * d + 2 * N_c * d + 2 * N_c
    per_c = cF + d + 2 * N_c * d + 3 * N_c
    per_face_node = cF + (8 + 4 * d) + 2 * N_c * d + 4 * N_c
    nv2 = 3 if ref.V is not None else 0
    f2 = (pairs * per_pair + nnzC * per_c + N_f * per_face_node + 2 * nnzR * N_c
          + nv2 * v * N_c + N_q * N_c)
    b2 = 8 * (N_q * N_c + 2 * N_f * N_c + N_q * d * d + N_q + d * N_f + N_f + N_p * N_c) + 4 * N_f
    logs = 2 * (pairs + nnzC + N_f) + N_q
    return {
        "nodal_values": KernelWork(float(f1), float(b1), float(N_q), float(N_q + 2 * (N_q + N_f))),
        "time_derivative": KernelWork(float(f2), float(b2), float(logs),
                                      float(14 * (pairs + nnzC + N_f))),
        "pairs": pairs, "nnzC": nnzC, "nnzR": nnzR,
    }
