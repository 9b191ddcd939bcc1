"""The proximity kernel's split scan (omega2d_amd/csrc/reflect.hip, near_kernel) as a CPU model:
four ranges of panels, each kept as an exact summary (its smallest distance m, its panels at m
and at next_up(m), whether one lies at next_up(next_up(m))), merged and replayed in index order
from the bound T = next_up(M) or next_up(next_up(M)). The model must end exactly where the
sequential scan of src/Reflect.h:490-826 ends (the final minimum and the final nearest set) on
orderings built from runs of consecutive doubles, where the scan's one-ulp windows make the
result depend on panel order."""
import numpy as np

DMAX = np.finfo(np.float64).max
KHITS = 4


def _bits(v):
    return int(np.array(v, dtype=np.float64).view(np.int64))


def _up(v):
    return v if v >= DMAX else np.nextafter(v, np.inf)


def _down(v):
    return np.nextafter(v, 0.0) if v > 0 else v


def sequential(ds, m0):
    """The reference's scan: below nextafter(m, 0) replaces the set, below nextafter(m, +max)
    joins it."""
    m, lo, hi, hits = m0, _down(m0), _up(m0), []
    for k, d in enumerate(ds):
        if d < lo:
            m, lo, hi, hits = d, _down(d), _up(d), [k]
        elif d < hi:
            hits.append(k)
    return m, hits


def split(ds, m0, waves=4):
    """The kernel's summaries per range, the choice of T and the replay; (result, replayed)."""
    per = (len(ds) + waves - 1) // waves
    summ = []
    for w in range(waves):
        bm, n0, n1, f2, l0, l1 = _bits(m0), 0, 0, 0, [], []
        for k in range(w * per, min(len(ds), (w + 1) * per)):
            dl = _bits(ds[k]) - bm
            if dl > 2:
                continue
            if dl < 0:
                if dl == -1:
                    f2, n1, l1 = int(n1 > 0), n0, l0
                else:
                    f2, n1, l1 = int(dl == -2), 0, []
                bm, n0, l0 = _bits(ds[k]), 1, [k]
            elif dl == 0:
                l0 = l0 + [k] if n0 < KHITS else l0
                n0 += 1
            elif dl == 1:
                l1 = l1 + [k] if n1 < KHITS else l1
                n1 += 1
            else:
                f2 = 1
        summ.append((bm, n0, n1, f2, l0, l1))
    M = min(s[0] for s in summ)
    at1 = any((s[0] == M and s[2] > 0) or s[0] == M + 1 for s in summ)
    at2 = any((s[0] == M and s[3]) or (s[0] == M + 1 and s[2] > 0) or s[0] == M + 2
              for s in summ)
    two = at1
    exact = (not at2 and M + 2 < _bits(m0)) if two else M + 1 < _bits(m0)
    below = []
    for bm, n0, n1, f2, l0, l1 in summ:
        if bm == M:
            exact &= n0 <= KHITS and (not two or n1 <= KHITS)
            below += l0 + (l1 if two else [])
        elif two and bm == M + 1:
            exact &= n0 <= KHITS
            below += l0
    if not exact:
        return sequential(ds, m0), False
    below.sort()
    m, hits = sequential([ds[k] for k in below], m0)
    return (m, [below[i] for i in hits]), True


def test_split_scan_ends_where_the_sequential_scan_ends():
    rng = np.random.default_rng(11)
    base = 0.7312
    replayed = 0
    for _ in range(20000):
        n = int(rng.integers(1, 41))
        ds = []
        for _ in range(n):
            if rng.random() < 0.6:
                v = base
                for _ in range(int(rng.integers(0, 5))):
                    v = np.nextafter(v, np.inf)
                ds.append(v)
            else:
                ds.append(base * (1.0 + rng.random()))
        m0 = DMAX if rng.random() < 0.5 else 0.1 * DMAX
        got, fast = split(ds, m0)
        ref = sequential(ds, m0)
        assert got[0] == ref[0] and got[1] == ref[1], ds
        replayed += fast
    assert replayed > 2000  # the replay path, not only the rescan, is exercised


def test_an_order_the_one_ulp_windows_decide():
    """next_up(M) before M: the sequential scan keeps next_up(M) as its minimum, and so must
    the split scan."""
    M = 0.09
    M1 = np.nextafter(M, np.inf)
    ds = [1.0] * 10 + [M1] + [1.0] * 10 + [M]
    assert sequential(ds, DMAX)[0] == M1
    got, fast = split(ds, DMAX)
    assert fast and got == sequential(ds, DMAX)
