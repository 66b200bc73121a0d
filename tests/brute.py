"""Independent brute-force restatements used to pin the oracle on tiny inputs.

Everything here is written from the mathematics, not from the oracle's code:
O(N^2) neighbor sets with an exactly rounded fmaf, exact integer sums, and
textbook closed forms.
"""
from fractions import Fraction

import numpy as np

F32 = np.float32
TWO24 = 16777216


def round_f32(fr):
    """Correctly rounds a Fraction to float32 (ties to even)."""
    c = np.float32(float(fr))
    best = None
    for cand in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        err = abs(Fraction(float(cand)) - fr)
        if best is None or err < best[0] or (
                err == best[0] and (int(np.float32(cand).view(np.uint32)) & 1) == 0):
            best = (err, cand)
    return np.float32(best[1])


def fmaf(a, b, c):
    """float32 fused multiply-add, exactly rounded."""
    return round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def d2_f32(qx, qy, px, py, tsx, tsy):
    tx = F32(F32(qx - px) * tsx)
    ty = F32(F32(qy - py) * tsy)
    return fmaf(tx, tx, F32(ty * ty))


def neighbor_mean_bruteforce(pts, fixed, res_x, res_y, radius):
    """All-pairs fixed-radius mean with exact fixed-point sums (the oracle's
    stated contract), O(N^2)."""
    pts = np.asarray(pts, dtype=np.float32)
    tsx, tsy = F32(F32(res_x) * F32(0.5)), F32(F32(res_y) * F32(0.5))
    r2 = F32(F32(radius) * F32(radius))
    qf = [(int(np.rint(F32(x) * F32(TWO24))), int(np.rint(F32(y) * F32(TWO24)))) for x, y in pts]
    out = pts.copy()
    for i, (px, py) in enumerate(pts):
        if fixed[i]:
            continue
        ax = ay = n = 0
        for j, (qx, qy) in enumerate(pts):
            tx = F32(F32(qx - px) * tsx)
            ty = F32(F32(qy - py) * tsy)
            approx = float(tx) * float(tx) + float(ty * ty)
            if abs(approx - float(r2)) > 1e-5 * float(r2):
                ok = approx <= float(r2)
            else:
                ok = fmaf(tx, tx, F32(ty * ty)) <= r2
            if ok:
                ax += qf[j][0]
                ay += qf[j][1]
                n += 1
        dx = ax - n * qf[i][0]
        dy = ay - n * qf[i][1]
        den = float(n) * float(TWO24)
        out[i, 0] = np.float32(float(px) + float(dx) / den)
        out[i, 1] = np.float32(float(py) + float(dy) / den)
    return out


def neighbor_mean_f64(pts, fixed, res_x, res_y, radius):
    """Plain float64 mean over the same neighbor sets (tolerance check)."""
    pts64 = np.asarray(pts, dtype=np.float64)
    out = np.asarray(pts, dtype=np.float64).copy()
    tsx, tsy = float(res_x) * 0.5, float(res_y) * 0.5
    for i in range(len(pts64)):
        if fixed[i]:
            continue
        d = (pts64 - pts64[i]) * [tsx, tsy]
        m = (d ** 2).sum(1) <= float(radius) ** 2
        out[i] = pts64[m].mean(0)
    return out
