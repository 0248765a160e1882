"""Closed-form cases for the reference's per-turn construct classes
(megno_construct, megno_birkhoff_construct, tune_birkhoff_construct;
src/henon.cu:2223-2655), shared by the oracle tests and the GPU tests.

Each case feeds a construct inputs whose result follows from the mathematics,
not from retyping the construct's recursion:

* MEGNO, scaled identities M_i = s_i I: every Frobenius ratio |M_i|/|M_{i-1}|
  is s_i/s_{i-1}.
  - constant ratio 2: L1_k = log10(2) k(k+1)/2, L2_k = sum_i L1_i/i, so
    megno_construct gives L3_k = log10(2) ((k+1)/4 + 1/2);
  - log10 ratio c/i ("harmonic" scales s_i = 10^(c H_i)): L1_k = c k, every
    L1_i/i = c, so L3_k = c for every k.
* megno_birkhoff_construct after at least T_j adds: L2_j = sum_{i<=T_j} L1_i
  w_j[i-1]/i. The Birkhoff weights w_j[m] = exp(-1/(t(1-t))), t = m/T_j, are
  normalised (sum 1) and symmetric (w[0] = 0, w[m] = w[T-m]), so
  sum_m m w[m] = T/2:
  - constant ratio 2: L2_j = (log10 2 / 2) sum_m (m+2) w[m] = log10(2) (1 + T_j/4);
  - harmonic: L2_j = c sum_m w[m] = c.
* tune_birkhoff_construct fed points whose phase angle falls by
  D_m = c0 + c1 m at the m-th add: the wrapped advance is 2 pi - D_m, so with
  the same two weight identities tune1_j = 2 pi - c0 - c1 T_j/2 (adds >= T_j - 1)
  and tune2_j = 2 pi - c0 - 3 c1 T_j / 2 (adds >= 2 T_j - 1).

A wrong weighting (L1 idx^2, a missing 1/idx), a shifted weight index (w[idx]
for w[idx-1] and back) or a wrong window start each moves these by O(1)."""
import math

import numpy as np

LOG2 = math.log10(2.0)
HARMONIC_C = 0.37


def megno_scales(kind, n_adds):
    """s_0 .. s_n_adds for the scaled identities M_i = s_i I."""
    i = np.arange(n_adds + 1, dtype=np.float64)
    if kind == "const2":
        return 2.0 ** i
    harm = np.concatenate([[0.0], np.cumsum(1.0 / np.arange(1, n_adds + 1))])
    return 10.0 ** (HARMONIC_C * harm)


def megno_expected(kind, k):
    """megno_construct's value (layer 3) after k adds."""
    if kind == "const2":
        return LOG2 * ((k + 1) / 4.0 + 0.5)
    return HARMONIC_C


def megno_birkhoff_expected(kind, t):
    """megno_birkhoff_construct's value for turn sample t after >= t adds."""
    if kind == "const2":
        return LOG2 * (1.0 + t / 4.0)
    return HARMONIC_C


def identities(s, n):
    """n copies of s I, shape (n, 4, 4)."""
    return np.tile(np.eye(4) * s, (n, 1, 1))


# phase decrement per add, D_m = c0 + c1 m, per plane
TUNE_X = (0.9, 1.3e-3)
TUNE_Y = (2.2, -7e-4)


def tune_points(n, n_adds, seed=3):
    """Points of n particles for first_add (index 0) and adds 1..n_adds: the
    (x, px) and (y, py) phase angles fall by D_m at add m, radii and starting
    angles differ per particle. Returns x, px, y, py of shape (n_adds+1, n)."""
    rng = np.random.default_rng(seed)
    rx, ry = rng.uniform(0.05, 0.5, n), rng.uniform(0.05, 0.5, n)
    th0x, th0y = rng.uniform(-np.pi, np.pi, n), rng.uniform(-np.pi, np.pi, n)
    m = np.arange(n_adds + 1, dtype=np.float64)

    def angles(c, th0):
        d = c[0] + c[1] * m
        d[0] = 0.0
        return th0[None, :] - np.cumsum(d)[:, None]

    tx, ty = angles(TUNE_X, th0x), angles(TUNE_Y, th0y)
    return rx * np.cos(tx), rx * np.sin(tx), ry * np.cos(ty), ry * np.sin(ty)


def tune_expected(c, t, which):
    """tune 1 (which=1) or tune 2 (which=2) of turn sample t."""
    return 2.0 * math.pi - c[0] - c[1] * t * (0.5 if which == 1 else 1.5)
