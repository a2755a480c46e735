"""The error model behind the margin band, measured on the matrix cores.

The band (csrc/kmc_api.hip, band_constants) bounds the MFMA value of x . w
(w = -2c) by two terms per precision form:
  operand rounding   |x~ . w~ - x . w| <= t_in ||x|| ||w||
  accumulation       |mfma(x~, w~) - x~ . w~| <= gA sum_k |x~_k w~_k|
with fp32: t_in = 0, gA = gamma_{D+2}; bf16: t_in = 2 ub + ub^2 (ub = 2^-8),
gA = gamma_{D+1}; fp8 e4m3 with per-row power-of-two scales: t_in = 2 u8 +
u8^2 + 2 (1 + u8) sqrt(D) eta + D eta^2 (u8 = 2^-4, eta = 2^-10 / 224),
gA = 2^-8 (gamma_m = m u / (1 - m u), u = 2^-24). kmc_probe_mfma runs the
distance kernel's MFMA chain on adversarial operands; both inequalities are
checked against FP64 for every one of the 32 x 32 products, D = 16 .. 256.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


def gam(m, u=U):
    return m * u / (1.0 - m * u)


def constants(prec, d):
    if prec == "fp32":
        return 0.0, gam(d + 2)
    if prec == "bf16":
        ub = 2.0 ** -8
        return 2 * ub + ub * ub, gam(d + 1)
    u8, eta = 2.0 ** -4, 2.0 ** -10 / 224.0
    t_in = 2 * u8 + u8 * u8 + 2 * (1 + u8) * math.sqrt(d) * eta + d * eta * eta
    return t_in, 2.0 ** -8


def operands(kind, d, seed):
    rng = np.random.default_rng(seed)
    sgn = lambda shape: np.where(rng.random(shape) < 0.5, -1.0, 1.0)  # noqa: E731
    if kind == "spread":      # exponents over 2^-20 .. 2^20 within a row
        A = sgn((32, d)) * rng.uniform(1, 2, (32, d)) * 2.0 ** rng.integers(-20, 21, (32, d))
        B = sgn((32, d)) * rng.uniform(1, 2, (32, d)) * 2.0 ** rng.integers(-20, 21, (32, d))
    elif kind == "cancel":    # the second half of B cancels the first exactly
        A = rng.standard_normal((32, d)) * 2.0 ** rng.integers(-6, 7, (32, 1))
        A[:, d // 2:] = A[:, :d // 2]
        B = rng.standard_normal((32, d))
        B[:, d // 2:] = -B[:, :d // 2]
    elif kind == "same_sign":  # no cancellation: the largest accumulated sum
        A = rng.uniform(1, 2, (32, d))
        B = rng.uniform(1, 2, (32, d))
    elif kind == "bimodal":   # 1e4 and 1e-4 entries side by side
        A = np.where(rng.random((32, d)) < 0.5, 1e4, 1e-4) * sgn((32, d))
        B = np.where(rng.random((32, d)) < 0.5, 1e4, 1e-4) * sgn((32, d))
    else:                     # blob-like: offsets plus noise, as distance operands
        A = -2.0 * (rng.uniform(-10, 10, (32, 1)) + rng.standard_normal((32, d)))
        B = rng.uniform(-10, 10, (32, 1)) + 2.0 * rng.standard_normal((32, d))
    return A.astype(np.float32), B.astype(np.float32)


KINDS = ["spread", "cancel", "same_sign", "bimodal", "blob"]


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_mfma_error_within_band_constants(kmc_mod, prec):
    dims = [64, 128, 256] if prec == "fp8" else [16, 32, 64, 128, 256]
    worst_acc = worst_op = 0.0
    for d in dims:
        t_in, gA = constants(prec, d)
        for kind in KINDS:
            for seed in range(3):
                A, B = operands(kind, d, 1000 * d + seed)
                out, Aq, Bq = kmc_mod.probe_mfma(prec, A, B)
                Aq64, Bq64 = Aq.astype(np.float64), Bq.astype(np.float64)
                exact_q = Aq64 @ Bq64.T
                abs_q = np.abs(Aq64) @ np.abs(Bq64).T
                acc_err = np.abs(out.astype(np.float64) - exact_q)
                ok = acc_err <= gA * abs_q
                assert ok.all(), (prec, d, kind, seed, float(np.max(acc_err - gA * abs_q)))
                nz = abs_q > 0
                worst_acc = max(worst_acc, float(np.max(acc_err[nz] / (gA * abs_q[nz]))))
                A64, B64 = A.astype(np.float64), B.astype(np.float64)
                op_err = np.abs(exact_q - A64 @ B64.T)
                norms = np.outer(np.linalg.norm(A64, axis=1), np.linalg.norm(B64, axis=1))
                assert np.all(op_err <= t_in * norms + 1e-300), (prec, d, kind, seed)
                if t_in > 0:
                    worst_op = max(worst_op, float(np.max(op_err / (t_in * norms))))
    print("%s: worst accumulation error / gA bound %.3g, worst operand error / t_in bound %.3g"
          % (prec, worst_acc, worst_op), flush=True)
    sys.stdout.flush()
