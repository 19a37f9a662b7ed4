"""Comparison of two text checkpoints at the precision they are printed
with (shared by the CPU and GPU checkpoint tests)."""
import numpy as np


def parse(text):
    """{ID: (particle fields, {partner: contact fields})}."""
    lines = text.split("\n")
    n = int(lines[0])
    out, pos = {}, 1
    for _ in range(n):
        tok = lines[pos].split()
        pos += 1
        cs = {}
        for _ in range(int(tok[-1])):
            ct = lines[pos].split()
            pos += 1
            cs[int(ct[0])] = np.array([float(t) for t in ct[1:]])
        out[int(tok[0])] = (np.array([float(t) for t in tok[1:-1]]), cs)
    return out


def close(a, b, rel):
    """Values printed with 9 significant digits from doubles that agree to
    `rel` relative differ by at most one unit of the ninth digit plus rel."""
    return np.all(np.abs(a - b) <= (1.01e-8 + rel) * np.abs(b) + 1e-300)


def compare(ours, ref, rel=0.0, scale_frac=0.0):
    """Assert two parsed checkpoints hold the same particles and contacts
    and agree at printed precision plus `rel` relative; scale_frac adds an
    absolute slack of that fraction of each contact field's largest value
    (for values that cancel to ~0)."""
    assert sorted(ours) == sorted(ref)
    rows = [v for p in ref for v in ref[p][1].values()]
    floor = (scale_frac * np.max(np.abs(np.array(rows)), axis=0) if rows else 0.0) + 1e-300
    for p in ref:
        pa, ca = ours[p]
        pb, cb = ref[p]
        assert close(pa, pb, rel), (p, pa, pb)
        assert sorted(ca) == sorted(cb), (p, sorted(ca), sorted(cb))
        for q in cb:
            a, b = ca[q], cb[q]
            ok = np.abs(a - b) <= (1.01e-8 + rel) * np.abs(b) + floor
            assert np.all(ok), (p, q, np.nonzero(~ok)[0], a[~ok], b[~ok])
