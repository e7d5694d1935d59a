"""Closed-form / brute-force statements of the reference semantics (numpy).

Used only by tests to pin the oracle: these are a third formulation, built
from set algebra rather than the reference's procedural sort / descent /
subcell lists that the oracle transcribes.
  - topology: the sorted node order, parents, child lists and atlas
    coordinates follow from the key definition of gvdbCalcBrickId
    (cuda_gvdb_particles.cuh:352-382), the extra-brick box of
    gvdbCalcExtraBrickId (467-512) and getAtlasPos (gvdb_allocator.cpp:705-715);
  - gathers: a voxel's subcell list is every point whose subcell box
    (cuda_gvdb_particles.cuh:237-242) contains the voxel's subcell;
  - apron: apron voxel = the world voxel's value in the leaf holding it.
"""
import numpy as np

NOHIT = np.float32(1.0e10)
UNDEFL = np.uint64(0xFFFFFFFF)


def ctrunc_div(a, b):
    """C integer division (truncation toward zero) for numpy int arrays."""
    a = np.asarray(a, dtype=np.int64)
    q = np.abs(a) // b
    return np.where(a < 0, -q, q)


def valid_points(points, orig):
    p = np.asarray(points, dtype=np.float32)
    o = np.asarray(orig, dtype=np.float32)
    fin = np.isfinite(p).all(1) & (p[:, 2] != NOHIT)
    w = (p + o).astype(np.float32)
    ok = fin & np.isfinite(w).all(1) & (w >= 0).all(1)
    return ok, w


def lexsort_cells(c):
    """Sort integer cell rows (x,y,z) with x most significant (reference key order)."""
    if len(c) == 0:
        return c.reshape(0, 3)
    order = np.lexsort((c[:, 2], c[:, 1], c[:, 0]))
    return c[order]


def unique_cells(c):
    if len(c) == 0:
        return c.reshape(0, 3)
    return lexsort_cells(np.unique(c, axis=0))


def expected_topology(points, radius, orig, log2dims, subdim_xy=(16, 16), apron=1):
    ok, w = valid_points(points, orig)
    w = w[ok]
    res = [1 << d for d in log2dims]
    rng = [int(np.prod(res[: l + 1])) for l in range(len(res))]
    mn, mx = w.min(0), w.max(0)
    pmin = np.maximum((mn - np.float32(1)).astype(np.float32), np.float32(0))
    pmax = np.maximum((mx + np.float32(2)).astype(np.float32), np.float32(0))
    ipmin, ipmax = pmin.astype(np.int64), pmax.astype(np.int64)
    depth = None
    for lev in range(len(res)):
        a0 = (ipmin // rng[lev]) * rng[lev]
        a1 = (ipmax // rng[lev]) * rng[lev]
        if (a0 == a1).all():
            depth, root = lev, a0
            break
    assert depth is not None and 0 < depth < len(res)
    base = [unique_cells(np.floor(w / np.float32(rng[l])).astype(np.int64)) for l in range(depth)]
    lists = []
    for l in range(depth):
        rg = rng[l]
        ext = np.zeros((0, 3), dtype=np.int64)
        if radius >= 1:
            ix = w.astype(np.int32)
            bmin = (ix.astype(np.float32) - np.float32(radius)).astype(np.int32).astype(np.int64)
            bmax = (ix.astype(np.float32) + np.float32(2) * np.float32(radius)).astype(np.int32).astype(np.int64)
            lo, hi = ctrunc_div(bmin, rg), ctrunc_div(bmax, rg)
            span = int((hi - lo).max()) + 1
            cells = []
            for dx in range(span):
                for dy in range(span):
                    for dz in range(span):
                        c = lo + np.array([dx, dy, dz])
                        m = (c <= hi).all(1)
                        cells.append(c[m])
            allc = np.unique(np.concatenate(cells), axis=0)
            inroot = ((allc * rg >= root) & (allc * rg < root + rng[depth])).all(1)
            allc = allc[inroot]
            bset = {tuple(x) for x in base[l]}
            ext = np.array([x for x in allc if tuple(x) not in bset], dtype=np.int64).reshape(-1, 3)
            ext = lexsort_cells(ext)
        lists.append(np.concatenate([base[l], ext]) if len(ext) else base[l])
    cnt = [len(lists[l]) for l in range(depth)] + [1] + [0] * (len(res) - depth - 1)
    top = 1
    for l in range(len(res) - 1, -1, -1):
        if cnt[l] == 1:
            top = l
    # node records
    nodes = []
    for l in range(depth):
        c = lists[l]
        n = len(c)
        rec = {
            "mLev": np.full(n, l), "mFlags": np.ones(n, dtype=np.int64), "mPos": c * rng[l],
            "mValue": np.full((n, 3), -1, dtype=np.int64),
            "mParent": np.full(n, UNDEFL, dtype=np.uint64),
            "mChildList": (np.full(n, UNDEFL, dtype=np.uint64) if l == 0 else
                           (np.uint64(1) | np.uint64(l << 8) | (np.arange(n, dtype=np.uint64) << np.uint64(16)))),
        }
        nodes.append(rec)
    nodes.append({"mLev": np.array([depth]), "mFlags": np.array([1]), "mPos": root.reshape(1, 3),
                  "mValue": np.full((1, 3), -1), "mParent": np.array([UNDEFL], dtype=np.uint64),
                  "mChildList": np.array([np.uint64(1) | np.uint64(depth << 8)], dtype=np.uint64)})
    clists = [None] + [np.full((cnt[l], res[l] ** 3), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
                       for l in range(1, depth + 1)]
    for l in range(depth):
        if l >= top:
            continue
        r = res[l + 1]
        pos = nodes[l]["mPos"]
        if l + 1 == depth:
            pidx = np.zeros(len(pos), dtype=np.int64)
            ppos = np.repeat(root.reshape(1, 3), len(pos), 0)
        else:
            pc = pos // rng[l + 1]
            lut = {tuple(x): i for i, x in enumerate(lists[l + 1])}
            pidx = np.array([lut[tuple(x)] for x in pc], dtype=np.int64)
            ppos = pc * rng[l + 1]
        loc = (pos - ppos) * r // rng[l + 1]
        bit = (loc[:, 2] * r + loc[:, 1]) * r + loc[:, 0]
        nodes[l]["mParent"] = (pidx.astype(np.uint64) << np.uint64(16)) | np.uint64((l + 1) << 8)
        clists[l + 1][pidx, bit] = (np.arange(len(pos), dtype=np.uint64) << np.uint64(16)) | np.uint64(l << 8)
    # atlas (getAtlasPos)
    sx, sy = subdim_xy
    br = res[0] + 2 * apron
    k = np.arange(cnt[0])
    z, rem = k // (sx * sy), k % (sx * sy)
    nodes[0]["mValue"] = np.stack([(rem % sx) * br + apron, (rem // sx) * br + apron, z * br + apron], 1)
    return {"depth": depth, "top_lev": top, "root": root, "cnt": cnt, "nodes": nodes, "clists": clists,
            "pos_min": pmin, "pos_max": pmax, "res": res, "range": rng}


def leaf_lookup(leaf_pos, res0):
    return {tuple((p // res0).tolist()): i for i, p in enumerate(leaf_pos)}


def subcell_box(w, R, sc):
    """Per-point voxel range [lo, hi] of the subcell box, C semantics."""
    lo = ctrunc_div((w - np.float32(R)).astype(np.float32).astype(np.int32), sc) * sc
    hi = ctrunc_div(((w + np.float32(1)).astype(np.float32) + np.float32(R)).astype(np.float32).astype(np.int32),
                    sc) * sc + sc - 1
    return lo, hi


def gather_f64(points, orig, leaf_pos, res0, sc, R, r, mode):
    """mode 1: min(3, min_members |p-v|-r); mode 0: max(0, max_members r-|p-v|). (B,res0,res0,res0)[z,y,x]."""
    ok, w = valid_points(points, orig)
    w = w[ok].astype(np.float64)
    lo, hi = subcell_box(w.astype(np.float32), R, sc)
    B = len(leaf_pos)
    out = np.zeros((B, res0, res0, res0))
    g = np.arange(res0)
    for k in range(B):
        p0 = leaf_pos[k]
        vz, vy, vx = np.meshgrid(g + p0[2], g + p0[1], g + p0[0], indexing="ij")
        v = np.stack([vx.ravel(), vy.ravel(), vz.ravel()], 1)
        s = v - v % sc
        near = ((hi >= p0) & (lo <= p0 + res0 - 1)).all(1)
        pw, plo, phi = w[near], lo[near], hi[near]
        if len(pw) == 0:
            val = np.full(len(v), 3.0 if mode == 1 else 0.0)
        else:
            mem = ((plo[None, :, :] <= s[:, None, :]) & (s[:, None, :] <= phi[None, :, :])).all(2)
            d = np.sqrt(((pw[None, :, :] - v[:, None, :]) ** 2).sum(2))
            if mode == 1:
                val = np.minimum(3.0, np.where(mem, d - r, np.inf).min(1))
            else:
                val = np.maximum(0.0, np.where(mem, r - d, -np.inf).max(1))
        out[k] = val.reshape(res0, res0, res0)
    return out


def cubic_f64(points, orig, leaf_pos, res0, r, weights=None):
    ok, w = valid_points(points, orig)
    wt = np.ones(len(points)) if weights is None else np.asarray(weights, dtype=np.float64)
    w, wt = w[ok].astype(np.float64), wt[ok]
    B = len(leaf_pos)
    out = np.zeros((B, res0, res0, res0))
    g = np.arange(res0)
    for k in range(B):
        p0 = leaf_pos[k]
        near = ((w >= p0 - r) & (w <= p0 + res0 - 1 + r)).all(1)
        if not near.any():
            continue
        vz, vy, vx = np.meshgrid(g + p0[2], g + p0[1], g + p0[0], indexing="ij")
        v = np.stack([vx.ravel(), vy.ravel(), vz.ravel()], 1)
        d2 = ((w[near][None, :, :] - v[:, None, :]) ** 2).sum(2)
        t = np.clip(1.0 - d2 / (r * r), 0.0, None)
        out[k] = (t ** 3 * wt[near][None, :]).sum(1).reshape(res0, res0, res0)
    return out


def apron_expected(bricks, leaf_pos, res0, boundval):
    """Recompute every apron voxel from the interiors (bricks: (B,br,br,br) [z,y,x]).
    bricks may be any 4-byte dtype (uint32 words of a colour channel: then
    boundval is the word stored outside every leaf)."""
    br = bricks.shape[1]
    a = (br - res0) // 2
    lut = leaf_lookup(leaf_pos, res0)
    out = bricks.copy()
    idx = np.arange(br)
    zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
    shell = ~((xx >= a) & (xx < a + res0) & (yy >= a) & (yy < a + res0) & (zz >= a) & (zz < a + res0))
    sz, sy, sx = zz[shell], yy[shell], xx[shell]
    for k in range(len(leaf_pos)):
        p0 = leaf_pos[k]
        u = np.stack([p0[0] + sx - a, p0[1] + sy - a, p0[2] + sz - a], 1)
        vals = np.full(len(u), boundval, dtype=bricks.dtype)
        for i, uu in enumerate(u):
            if (uu < 0).any():
                continue
            j = lut.get(tuple((uu // res0).tolist()))
            if j is not None:
                l = uu - leaf_pos[j] + a
                vals[i] = bricks[j, l[2], l[1], l[0]]
        out[k][shell] = vals
    return out


def subcell_members(points, orig, leaf_pos, res0, sc, R):
    """Brute-force subcell lists: for leaf k and subcell s (x fastest), the
    indices of the valid points whose insertion box holds the subcell origin."""
    ok, w = valid_points(points, orig)
    idx = np.nonzero(ok)[0]
    w = w[ok].astype(np.float32)
    lo, hi = subcell_box(w, R, sc)
    nsd = res0 // sc
    out = []
    g = np.arange(nsd) * sc
    oz, oy, ox = np.meshgrid(g, g, g, indexing="ij")
    offs = np.stack([ox.ravel(), oy.ravel(), oz.ravel()], 1)
    for k in range(len(leaf_pos)):
        o = leaf_pos[k][None, :] + offs
        mem = ((lo[None, :, :] <= o[:, None, :]) & (o[:, None, :] <= hi[None, :, :])).all(2)
        out.append([idx[np.nonzero(mem[s])[0]] for s in range(len(offs))])
    return out


def sat_u8(x):
    """float32 -> uint8 as a saturating truncation (NaN -> 0)."""
    x = np.asarray(x, dtype=np.float32)
    with np.errstate(invalid="ignore"):
        y = np.where(np.isnan(x) | (x <= 0), 0, np.where(x >= 255, 255, np.trunc(np.nan_to_num(x))))
    return y.astype(np.uint8)


def gather_color_expected(members, clr, res0, sc, init=None):
    """Colour mean per interior voxel (B, res0, res0, res0, 4) [z,y,x,c] from
    brute-force lists: f = init + float32(S)/255, sat(trunc(f/n * 255)) with n
    = count (+1 when init, the accumulated colour, is given)."""
    nsd = res0 // sc
    B = len(members)
    out = np.zeros((B, res0, res0, res0, 4), np.uint8)
    c64 = clr.astype(np.int64)
    for k in range(B):
        for s, m in enumerate(members[k]):
            sx, sy, sz = s % nsd, (s // nsd) % nsd, s // (nsd * nsd)
            S = c64[m].sum(0) if len(m) else np.zeros(4, np.int64)
            sl = (k, slice(sz * sc, sz * sc + sc), slice(sy * sc, sy * sc + sc), slice(sx * sc, sx * sc + sc))
            if init is None:
                base = np.broadcast_to(np.array([0, 0, 0, 1], np.float32), (sc, sc, sc, 4))
                n = np.float32(len(m))
            else:
                base = init[sl].astype(np.float32) / np.float32(255)
                n = np.float32(len(m) + 1)
            with np.errstate(divide="ignore", invalid="ignore"):
                f = base + S.astype(np.float32) / np.float32(255)
                out[sl] = sat_u8((f / n) * np.float32(255))
    return out


def incre_touched(points, radius, orig, rng, depth, root):
    """Cells of every level below the root that the incremental update's box
    [int(w) - r, int(w) + r] reaches (gvdbCalcIncreExtraBrickId,
    cuda_gvdb_particles.cuh:420-465), C truncation of the corner division,
    cells outside the root dropped; per level, sorted in key order."""
    ok, w = valid_points(points, orig)
    w = w[ok]
    ix = w.astype(np.int32).astype(np.float32)
    bmin = (ix - np.float32(radius)).astype(np.int32).astype(np.int64)
    bmax = (ix + np.float32(radius)).astype(np.int32).astype(np.int64)
    root = np.asarray(root, dtype=np.int64)
    out = []
    for l in range(depth):
        rg = rng[l]
        lo, hi = ctrunc_div(bmin, rg), ctrunc_div(bmax, rg)
        span = int((hi - lo).max()) + 1 if len(lo) else 0
        cells = [np.zeros((0, 3), dtype=np.int64)]
        for dx in range(span):
            for dy in range(span):
                for dz in range(span):
                    c = lo + np.array([dx, dy, dz])
                    cells.append(c[(c <= hi).all(1)])
        allc = np.unique(np.concatenate(cells), axis=0)
        inroot = ((allc * rg >= root) & (allc * rg < root + rng[depth])).all(1)
        out.append(lexsort_cells(allc[inroot]))
    return out


def reachable(vol, depth):
    """Per level below the root, the node indices reached from the root
    through the child lists (entries (index << 16) | (level << 8))."""
    res_lists = {l: vol.child_lists(l) for l in range(1, depth + 1)}
    nodes = {l: vol.nodes(l) for l in range(depth + 1)}
    reach = {depth: np.array([0])}
    for l in range(depth, 0, -1):
        cl = nodes[l]["mChildList"][reach[l]] >> np.uint64(16)
        ent = res_lists[l][cl.astype(np.int64)].ravel()
        ent = ent[ent != np.uint64(0xFFFFFFFFFFFFFFFF)]
        assert ((ent >> np.uint64(8)) & np.uint64(0xFF) == np.uint64(l - 1)).all()
        reach[l - 1] = np.sort((ent >> np.uint64(16)).astype(np.int64))
    return reach


def reachable_cells(vol, depth, rng):
    """reachable() as sorted cell rows per level."""
    reach = reachable(vol, depth)
    return [lexsort_cells(vol.nodes(l)["mPos"][reach[l]].astype(np.int64) // rng[l]) for l in range(depth)]
