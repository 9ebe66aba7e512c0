"""Host-side layout builders, checked on the CPU against brute force.

tests/host_layout_dump.cpp links sors_amd/csrc/pattern.cpp and coarse.cpp,
runs build_pattern and build_coarse on a mesh and dumps their arrays.  The
GPU kernels trust these arrays blindly, so each is checked here from its
definition:

- slot -> element-block lists (two-phase assembly): every kept element block
  (e, la, lb) appears exactly once, in the list of the slot of node conn[e,la]
  whose column is conn[e,lb], lists in ascending element order (tet4 entries
  name one of the element's ten stored upper-triangle blocks and whether it is
  read transposed);
- node -> incident (element, local node) lists;
- aggregates: every free node in exactly one, Dirichlet nodes in none;
  centroids, RMS radii, the rotation flag and the FP32 rigid-body offsets;
- Galerkin pair lists: every fine block between free nodes exactly once,
  under the pair (agg(i), agg(j)), with its value index, chunked by <= 256.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sors_amd import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sors_amd", "csrc")
SKIP = 255


@pytest.fixture(scope="module")
def dump_tool(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("hostlayout") / "host_layout_dump")
    subprocess.run([cxx, "-std=c++17", "-O1", "-pthread", "-I", CSRC, os.path.join(ROOT, "tests", "host_layout_dump.cpp"),
                    os.path.join(CSRC, "pattern.cpp"), os.path.join(CSRC, "coarse.cpp"), "-o", out], check=True)
    return out


def run_dump(tool, tmp_path, X, conn, mask, target):
    N, E, ed = X.shape[0], conn.shape[0], conn.shape[1]
    src, dst = tmp_path / "mesh.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([N, E, ed, target], dtype=np.int32).tofile(f)
        np.ascontiguousarray(conn, dtype=np.int32).tofile(f)
        np.ascontiguousarray(X, dtype=np.float64).reshape(-1).tofile(f)
        np.ascontiguousarray(mask, dtype=np.int32).tofile(f)
    subprocess.run([tool, str(src), str(dst)], check=True)
    out = {}
    raw = dst.read_bytes()
    off = 0
    while off < len(raw):
        name = raw[off:off + 32].split(b"\0")[0].decode()
        es = int(np.frombuffer(raw, np.int32, 1, off + 32)[0])
        n = int(np.frombuffer(raw, np.int64, 1, off + 36)[0])
        off += 44
        dt = {("rank", 1): np.uint8, ("cent", 8): np.float64, ("scale", 8): np.float64, ("dn", 4): np.float32}.get(
            (name, es), np.int32)
        out[name] = np.frombuffer(raw, dt, n, off).copy()
        off += n * es
    return out


def cases():
    t = mesh.beam(nx=9, ny=3, nz=3, element="tet", e_range=(5e5, 5e6), seed=1)
    h = mesh.beam(nx=6, ny=3, nz=2, element="hex", e_range=(5e5, 5e6), seed=1)
    return [("tet_clamped", t.X, t.conn, t.bcmask, 16), ("hex_clamped", h.X, h.conn, h.bcmask, 20),
            ("tet_free_coarsest", t.X, t.conn, np.zeros_like(t.bcmask), 100000)]


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_slot_lists_and_incidence(dump_tool, tmp_path, case):
    _, X, conn, mask, target = case
    X = np.asarray(X).reshape(-1, 3)
    conn = np.asarray(conn)
    N, (E, ed) = X.shape[0], conn.shape
    d = run_dump(dump_tool, tmp_path, X, conn, np.asarray(mask, dtype=np.int32), target)
    sbase, bcol, rank = d["sbase"], d["bcol"], d["rank"].reshape(ed * ed, E)
    S = int(sbase[-1])
    nodefree = np.asarray(mask).reshape(-1, 3)[:, 0] == 0
    # expected (slot position, block id) for every kept element block
    e_idx = np.repeat(np.arange(E), ed * ed)
    la = np.tile(np.repeat(np.arange(ed), ed), E)
    lb = np.tile(np.tile(np.arange(ed), ed), E)
    r = rank[la * ed + lb, e_idx].astype(np.int64)
    v, w = conn[e_idx, la], conn[e_idx, lb]
    kept = r != SKIP
    # a pair is kept exactly when both nodes are free
    assert np.array_equal(kept, nodefree[v] & nodefree[w])
    p = sbase[v >> 6].astype(np.int64) + 64 * r + (v & 63)
    if ed == 4:
        # tet4: the ten stored blocks a <= b of the symmetric element Hessian,
        # row-major; block (la, lb) with la > lb is stored block (lb, la), read transposed
        a_, b_ = np.minimum(la, lb), np.maximum(la, lb)
        t = 4 * a_ - a_ * (a_ - 1) // 2 + b_ - a_
        upper = [(x, y) for x in range(4) for y in range(x, 4)]
        assert [4 * x - x * (x - 1) // 2 + y - x for x, y in upper] == list(range(10))
        # an element's ten stored blocks are contiguous: block t of element e at 10 e + t
        code = 2 * (10 * e_idx + t) + (la > lb)
    else:
        # hex8: all 64 blocks, tiled by 64 elements: block 8 la + lb of element e
        code = ((e_idx >> 6) * ed * ed + ed * la + lb) * 64 + (e_idx & 63)
    p, bid = p[kept], code[kept]
    assert np.array_equal(bcol[p], w[kept])  # the slot's column is the block's column node
    order = np.argsort(p, kind="stable")  # block ids ascend within each slot
    counts = np.bincount(p, minlength=S)
    assert np.array_equal(d["sptr"], np.concatenate([[0], np.cumsum(counts)]))
    assert np.array_equal(d["sidx"], bid[order])
    # each slot position's 64-group belongs to its slice
    q = np.arange(S // 64)
    assert np.array_equal(d["qslice"], np.searchsorted(sbase, 64 * q, side="right") - 1)
    # incidence: node -> e * ed + a, element order
    inc_exp = [[] for _ in range(N)]
    for e in range(E):
        for a in range(ed):
            inc_exp[conn[e, a]].append(e * ed + a)
    incptr, inc = d["incptr"], d["inc"]
    for node in range(N):
        assert list(inc[incptr[node]:incptr[node + 1]]) == inc_exp[node]


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_coarse_space(dump_tool, tmp_path, case):
    _, X, conn, mask, target = case
    X = np.asarray(X).reshape(-1, 3)
    conn = np.asarray(conn)
    N = X.shape[0]
    mask = np.asarray(mask, dtype=np.int32)
    d = run_dump(dump_tool, tmp_path, X, conn, mask, target)
    agg, aptr, anodes = d["agg"], d["aptr"], d["anodes"]
    nagg = len(aptr) - 1
    free = mask.reshape(-1, 3)[:, 0] == 0
    assert np.array_equal(agg < 0, ~free)
    assert agg.max() == nagg - 1 and len(anodes) == free.sum()
    cent, scale, rot = d["cent"].reshape(-1, 3), d["scale"], d["rot"]
    dn = d["dn"].reshape(N, 4)
    assert np.array_equal(dn[:, 3].view(np.int32), agg)
    for a in range(nagg):
        nodes = anodes[aptr[a]:aptr[a + 1]]
        assert np.array_equal(nodes, np.flatnonzero(agg == a))  # ascending node order
        c = X[nodes].mean(axis=0)
        dv = X[nodes] - c
        s = np.sqrt((dv * dv).sum() / len(nodes))
        assert np.allclose(cent[a], c, rtol=0, atol=1e-14) and abs(scale[a] - s) <= 1e-14 * max(s, 1)
        Sm = dv.T @ dv
        m2 = Sm[0, 0] * Sm[1, 1] - Sm[0, 1] ** 2 + Sm[0, 0] * Sm[2, 2] - Sm[0, 2] ** 2 + Sm[1, 1] * Sm[2, 2] - Sm[1, 2] ** 2
        tr = np.trace(Sm)
        assert rot[a] == int(len(nodes) >= 3 and tr > 0 and m2 > 1e-6 * tr * tr)
        exp = (dv / s).astype(np.float32) if rot[a] else np.zeros_like(dv, dtype=np.float32)
        assert np.allclose(dn[nodes, :3], exp, rtol=2e-7, atol=1e-7)
    if "coarsest" in case[0]:
        # target above the node count: cubic bins of the bounding box's
        # volume, so a 3:1:1 beam still gets ceil(3 / 3^(1/3)) = 3 of them
        assert nagg == 3
    else:
        assert nagg > 3
    # Galerkin pair lists: every fine block between free nodes, once
    sbase, bcol, self_slot = d["sbase"], d["bcol"], d["self_slot"]
    exp = set()
    for sl in range(len(sbase) - 1):
        width = (sbase[sl + 1] - sbase[sl]) // 64
        for lane in range(64):
            i = sl * 64 + lane
            if i >= N or not free[i]:
                continue
            for k in range(width):
                j = int(bcol[sbase[sl] + 64 * k + lane])
                if not free[j] or (j == i and k != self_slot[i]):
                    continue
                exp.add((i, j, 9 * int(sbase[sl]) + 64 * 9 * k + lane))
    pairI, pairJ, pptr = d["pairI"], d["pairJ"], d["pptr"]
    ci, cj, cbase = d["ci"], d["cj"], d["cbase"]
    got = list(zip(ci.tolist(), cj.tolist(), cbase.tolist()))
    assert len(got) == len(exp) and set(got) == exp
    assert np.all(np.diff(pairI) >= 0)
    for pidx in range(len(pairI)):
        lo, hi = pptr[pidx], pptr[pidx + 1]
        assert hi > lo
        assert np.all(agg[ci[lo:hi]] == pairI[pidx]) and np.all(agg[cj[lo:hi]] == pairJ[pidx])
        if pidx + 1 < len(pairI) and pairI[pidx + 1] == pairI[pidx]:
            assert pairJ[pidx + 1] > pairJ[pidx]
    # chunks of at most 256 contributions, inside their pair
    gptr, pchunk = d["gptr"], d["pchunk"]
    assert gptr[0] == 0 and gptr[-1] == len(ci) and np.all(np.diff(gptr) > 0) and np.all(np.diff(gptr) <= 256)
    for pidx in range(len(pairI)):
        ch = gptr[pchunk[pidx]:pchunk[pidx + 1] + 1]
        assert ch[0] == pptr[pidx] and ch[-1] == pptr[pidx + 1]
