"""Host-side logic that needs no GPU: synthetic workloads, the ctypes mirror of the C ABI."""
import math
import os
import re

import numpy as np

from ibhip import _lib
from ibhip.workloads import ellipse, ellipse_batch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ellipse_points_on_curve_and_spaced_by_h():
    h = 0.02
    b = ellipse((0.3, -0.2), 0.5, 0.2, 0.7, h)
    x, y = b.xb[:, 0] - 0.3, b.xb[:, 1] + 0.2
    c, s = math.cos(0.7), math.sin(0.7)
    u, v = c * x + s * y, -s * x + c * y  # back to the ellipse frame
    assert np.allclose((u / 0.5) ** 2 + (v / 0.2) ** 2, 1.0, atol=1e-12)
    seg = np.hypot(np.diff(np.r_[b.xb[:, 0], b.xb[0, 0]]), np.diff(np.r_[b.xb[:, 1], b.xb[0, 1]]))
    assert np.allclose(seg, h, rtol=0.05)
    assert np.allclose(b.ds, b.ds[0]) and abs(b.ds.sum() - seg.sum()) < 1e-3 * seg.sum()


def test_ellipse_batch_is_the_seeded_sweep():
    probs = ellipse_batch()
    assert len(probs) == 64
    again = ellipse_batch()
    for p, q in zip(probs, again):
        assert p.fluid.nx == q.fluid.nx and p.fluid.Re == q.fluid.Re
        assert np.array_equal(p.bodies[0].xb, q.bodies[0].xb)
    for p in probs:
        n, L = p.fluid.nx, p.fluid.nlevels
        assert n in (128, 192, 256, 384, 512) and p.fluid.ny == n and 2 <= L <= 5
        h = p.fluid.domain.base.h
        assert abs(h - 4.0 / n) < 1e-15 and abs(p.scheme.dt - 0.2 * h) < 1e-15
        assert 50 <= p.fluid.Re <= 400
        assert p.scheme.precision == "fp32" and p.scheme.force_solver == "dense"
        # the 13×13 delta support (±6h) stays inside the finest grid [-1, 3] × [-2, 2]
        xb = p.bodies[0].xb
        assert xb[:, 0].min() - 7 * h > -1.0 and xb[:, 0].max() + 7 * h < 3.0
        assert xb[:, 1].min() - 7 * h > -2.0 and xb[:, 1].max() + 7 * h < 2.0
    sizes = {p.fluid.nx for p in probs}
    levels = {p.fluid.nlevels for p in probs}
    assert len(sizes) == 5 and len(levels) == 4


def test_case_desc_mirrors_the_c_struct():
    # field names and order of ibh_case_desc in include/ibh.h against the ctypes Structure
    src = open(os.path.join(ROOT, "include", "ibh.h")).read()
    body = re.search(r"typedef struct ibh_case_desc \{(.*?)\} ibh_case_desc;", src, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        decl = re.sub(r"^(const\s+)?(int32_t|double)\s*\**", "", decl)
        for part in decl.split(","):
            names.append(re.sub(r"[\s*]|\[\d+\]", "", part))
    assert names == [f[0] for f in _lib.CaseDesc._fields_]
