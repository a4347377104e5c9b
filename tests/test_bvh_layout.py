"""The traversal layout of a build tree: depth-first f32 binary nodes, 4-wide
q8 nodes with byte-quantised child boxes, and the leaf triangle order.

CPU: the oracle's layout (oracle/oracle_layout.cpp) is pinned on trees the
oracle builds: walking the f32 tree next to the build tree meets every
interior node once with its children's exact boxes; the q8 tree reaches every
triangle exactly once, its decoded child boxes (float32 origin + q * 2^e)
contain the true boxes, the chosen step is the smallest power of two that
does, and its children are the collapse of the binary tree (the largest-area
interior child opened first); leaves come in depth-first order with
ascending primitive ids and a flag on the last. A planted slot or byte error
fails these walks.
GPU: the device layout passes (bvh_layout.hip) produce the same arrays, bit
for bit, from the build tree the device built."""
import numpy as np
import pytest

from lighthouse2_amd import scenes
import oracle as orc

REF_LEAF = 0x80000000


def f32v(u):
    return np.asarray(u, np.uint32).view(np.float32)


def kids(nodes, x):
    L = int(nodes[x]["leftFirst"])
    return L, L + 1


def is_leaf(nodes, x):
    return int(nodes[x]["count"]) > 0


def dfs_leaves(nodes):
    out, st = [], [0]
    while st:
        x = st.pop()
        if is_leaf(nodes, x):
            out.append(x)
            continue
        L, R = kids(nodes, x)
        st += [R, L]
    return out


def leaf_slots(nodes):
    base, slots = 0, {}
    for x in dfs_leaves(nodes):
        slots[x] = base
        base += int(nodes[x]["count"])
    return slots


def check_order(nodes, prims, order, last):
    exp, exp_last = [], []
    for x in dfs_leaves(nodes):
        f, c = int(nodes[x]["leftFirst"]), int(nodes[x]["count"])
        exp += sorted(int(p) for p in prims[f:f + c])
        exp_last += [0] * (c - 1) + [1]
    assert np.array_equal(order, np.array(exp, np.uint32))
    assert np.array_equal(last, np.array(exp_last, np.uint8))
    assert np.array_equal(np.sort(order), np.arange(len(prims), dtype=np.uint32))


def check_f32(nodes, f32):
    slots = leaf_slots(nodes)
    seen = set()
    st = [] if is_leaf(nodes, 0) else [(0, 0)]
    while st:
        x, s = st.pop()
        assert s not in seen
        seen.add(s)
        row = f32[s]
        L, R = kids(nodes, x)
        box = f32v(row[:12])
        nl, nr = nodes[L], nodes[R]
        assert np.array_equal(box[0:4], np.float32([nl["lo"][0], nl["hi"][0], nl["lo"][1], nl["hi"][1]]))
        assert np.array_equal(box[4:8], np.float32([nr["lo"][0], nr["hi"][0], nr["lo"][1], nr["hi"][1]]))
        assert np.array_equal(box[8:12], np.float32([nl["lo"][2], nl["hi"][2], nr["lo"][2], nr["hi"][2]]))
        for c, ref in ((L, int(row[12])), (R, int(row[13]))):
            if is_leaf(nodes, c):
                assert ref == REF_LEAF | slots[c]
            else:
                assert ref & REF_LEAF == 0 and ref > s
                st.append((c, ref))
    assert len(seen) == len(f32)


Q8_EMPTY = 0xFFFFFFFF


def decode_q8(row):
    """origin.xyz, plane steps as floats whose exponent fields are ex, ey, ez
    (qb.z, qb.w, origin.w), the child count from the unused slots' refs."""
    origin = f32v(row[0:3])
    steps = [int(row[10]), int(row[11]), int(row[3])]
    assert all(b & 0x807FFFFF == 0 for b in steps)          # exactly 2^(e - 127), 0 or inf
    ex = [b >> 23 for b in steps]
    refs = [int(r) for r in row[12:16]]
    cnt = sum(r != Q8_EMPTY for r in refs)
    assert all(r != Q8_EMPTY for r in refs[:cnt]) and all(r == Q8_EMPTY for r in refs[cnt:])
    q = np.array([[(int(row[4 + k]) >> (8 * c)) & 255 for c in range(4)] for k in range(6)])   # lo.xyz, hi.xyz
    return origin, ex, cnt, q, refs


def contains(lo, hi, origin, e, qlo, qhi):
    step = np.float32(2.0 ** (e - 127)) if e < 255 else np.float32(np.inf)
    dlo = np.float32(origin + np.float32(np.float32(qlo) * step))
    dhi = np.float32(origin + np.float32(np.float32(qhi) * step))
    return bool(dlo <= lo) and bool(dhi >= hi)


def outward_bytes(lo_child, hi_child, origin, e):
    """Bytes of one child on one axis rounded outward at exponent e (None if 255 steps do not reach)."""
    step = np.float32(2.0 ** (e - 127))
    a = int(max(0, min(255, np.floor((float(lo_child) - float(origin)) / float(step)))))
    while a > 0 and np.float32(origin + np.float32(np.float32(a) * step)) > lo_child:
        a -= 1
    b = int(max(0, min(255, np.ceil((float(hi_child) - float(origin)) / float(step)))))
    while b < 255 and np.float32(origin + np.float32(np.float32(b) * step)) < hi_child:
        b += 1
    ok = contains(lo_child, hi_child, origin, e, a, b)
    return (a, b) if ok else None


def half_area(nodes, x):
    d = (nodes[x]["hi"] - nodes[x]["lo"]).astype(np.float32)
    return np.float32(np.float32(np.float32(d[0] * d[1]) + np.float32(d[1] * d[2])) + np.float32(d[2] * d[0]))


def collapse(nodes, x):
    k = list(kids(nodes, x))
    while len(k) < 4:
        cand = [(half_area(nodes, c), -j) for j, c in enumerate(k) if not is_leaf(nodes, c)]
        if not cand:
            break
        j = -max(cand)[1]
        c = k[j]
        k[j], _ = kids(nodes, c)
        k.append(kids(nodes, c)[1])
    return k


def check_q8(nodes, q8, q8map, last, check_minimal=True):
    slots = leaf_slots(nodes)
    reached = np.zeros(len(last), np.int32)
    seen = set()
    st = [(0, 0)]
    if is_leaf(nodes, 0):                              # a root leaf: no q8 node, every triangle below it
        st, reached[:] = [], 1
    while st:
        x, s = st.pop()
        assert s not in seen
        seen.add(s)
        origin, ex, cnt, q, refs = decode_q8(q8[s])
        ch = collapse(nodes, x)
        assert cnt == len(ch) and list(q8map[s][:cnt + 1]) == [x] + ch
        for c, y in enumerate(ch):
            ref = refs[c]
            if is_leaf(nodes, y):
                assert ref == REF_LEAF | slots[y]
                k = slots[y]
                while True:
                    reached[k] += 1
                    if last[k]:
                        break
                    k += 1
            else:
                assert ref & REF_LEAF == 0 and ref > s
                st.append((y, ref))
            for a in range(3):
                assert contains(nodes[y]["lo"][a], nodes[y]["hi"][a], origin[a], ex[a], q[a][c], q[3 + a][c])
        if check_minimal:
            for a in range(3):
                ext = float(np.float32(nodes[x]["hi"][a] - nodes[x]["lo"][a]))
                floor_e = max(-126, int(np.ceil(np.log2(ext / 255.0)))) + 127 if ext > 0 else 1
                if ex[a] > floor_e:
                    e = ex[a] - 1
                    assert any(outward_bytes(nodes[y]["lo"][a], nodes[y]["hi"][a], origin[a], e) is None for y in ch)
    assert len(seen) == len(q8)
    assert np.all(reached == 1)


def duplicate_scene(copies):
    """Coincident centroids: leaves beyond the SAH leaf size."""
    v = np.tile(scenes.random_triangle_scene(40, seed=2).meshes[0][0][:3], (copies, 1))
    s = scenes.Scene()
    s.add_mesh(v, None)
    s.instances.append((0, np.eye(4, dtype=np.float32).reshape(16)))
    return s


def oracle_trees(s):
    o = orc.OracleScene(s)
    for m in range(len(s.meshes)):
        nodes, prims, _ = o.bvh_nodes(m)
        yield nodes, prims


@pytest.mark.parametrize("scene", ["random", "cornell", "duplicates"])
def test_oracle_layout_invariants(scene):
    s = {"random": lambda: scenes.random_triangle_scene(6000, seed=11),
         "cornell": lambda: scenes.cornell_box(sphere_tris=2000),
         "duplicates": lambda: duplicate_scene(600)}[scene]()
    for nodes, prims in oracle_trees(s):
        f32, q8, order, last, counts, q8map = orc.bvh_layout(nodes, prims)
        check_order(nodes, prims, order, last)
        check_f32(nodes, f32)
        check_q8(nodes, q8, q8map, last)
        assert counts[0] == (len(nodes) - 1) // 2 and counts[4] >= 1


def test_planted_layout_errors_fail_the_walks():
    s = scenes.random_triangle_scene(2000, seed=4)
    nodes, prims = next(oracle_trees(s))
    f32, q8, order, last, counts, q8map = orc.bvh_layout(nodes, prims)
    bad = f32.copy()
    i = next(i for i in range(len(bad)) if bad[i][13] & REF_LEAF == 0)
    bad[i][13] += 1                                    # right child slot off by one
    with pytest.raises(AssertionError):
        check_f32(nodes, bad)
    bad = q8.copy()
    bad[0][4] = (int(bad[0][4]) & ~0xFF) | min(255, (int(bad[0][4]) & 0xFF) + 1)  # child 0's lo.x one step inward
    with pytest.raises(AssertionError):
        check_q8(nodes, bad, q8map, last)
    bad = q8.copy()
    w = int(bad[0][3])
    bad[0][3] = (w & ~0xFF) | ((w & 0xFF) + 1)         # a coarser step than needed on x
    with pytest.raises(AssertionError):
        check_q8(nodes, bad, q8map, last)
    bad_order = order.copy()
    bad_order[[0, 1]] = bad_order[[1, 0]]
    with pytest.raises(AssertionError):
        check_order(nodes, prims, bad_order, last)


@pytest.mark.gpu
@pytest.mark.parametrize("scene", ["random", "cornell", "duplicates", "large"])
def test_device_layout_equals_oracle_layout(torch_cuda, scene):
    import lighthouse2_amd as lh
    s = {"random": lambda: scenes.random_triangle_scene(30000, seed=11),
         "large": lambda: scenes.random_triangle_scene(300000, seed=12),
         "cornell": lambda: scenes.cornell_box(),
         "duplicates": lambda: duplicate_scene(600)}[scene]()
    print(f"device layout: scene {scene}", flush=True)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    for m in range(len(s.meshes)):
        nodes, prims = core.bvh_download(m)
        f32, q8, order, last, counts, q8map = orc.bvh_layout(nodes, prims)
        gf32, gq8, gtris = core.bvh_layout_download(m)
        info = core.bvh_info(m)
        assert info["nodes_f32"] == counts[0] and info["nodes_q8"] == counts[1] and info["depth_q8"] == counts[5]
        assert np.array_equal(gf32, f32) and np.array_equal(gq8, q8)
        w0 = gtris[:, 3].view(np.uint32)
        assert np.array_equal(w0 & 0x7FFFFFFF, order) and np.array_equal((w0 >> 31).astype(np.uint8), last)
        exp = s.meshes[m][0].reshape(-1, 3, 4)[order].reshape(-1, 12).copy()
        exp[:, 3] = gtris[:, 3]
        assert np.array_equal(gtris, exp)
    core.Shutdown()
