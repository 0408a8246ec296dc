"""Pins of the oracle's voxel hair technique (GenerateVoxelMesh, geometry_processor.cpp:515-602,
and the DDA walk that replaces the placeholder voxel_intersection.rint):

- the grid: NextDivisible and the truncated resolution in float32 (numpy's float32 is the same
  IEEE arithmetic);
- an x-major axis-aligned line with dyadic coordinates (every float step exact): the full sweep
  fills exactly the box of cells, the reference sweep exactly the cells at each step's x;
- lines in a coordinate plane (where the reference's projected radii are the exact slice
  extents): every point of the capsule lies in a filled cell (full sweep), and no filled cell
  is farther from the segment than the sweep can reach;
- the reference sweep on a y-major line at constant x fills the box-minimum x column only (its
  fill loop fixes x), the full sweep both columns the radius reaches;
- each level of detail is the OR of its 2x2x2 children;
- the walk against a brute-force slab test over every filled cell, and closed forms for rays
  along an axis;
- the device walk's block moves against the cell walk (host C++ restatement of both)."""
import numpy as np
import pytest


def _bits_to_cells(bits, res):
    n = int(res[0]) * int(res[1]) * int(res[2])
    flat = np.unpackbits(bits.view(np.uint8), bitorder="little")[:n].astype(bool)
    return flat.reshape(int(res[2]), int(res[1]), int(res[0]))  # [z, y, x]


def test_grid_parameters(oracle):
    f = np.float32
    for lo, hi, s in [((0, 0, 0), (1, 0.5, 0.25), 0.25), ((-2.3, 1.1, 0.0), (4.7, 1.3, 9.99), 0.1),
                      ((0.1, 0.2, 0.3), (0.4, 0.5, 0.6), 0.1), ((5, 5, 5), (5, 6, 7), 0.3)]:
        lo32, hi32, s32 = np.array(lo, f), np.array(hi, f), f(s)
        lines = np.zeros((1, 8), f)
        lines[0, :3], lines[0, 4:7] = lo32, lo32
        _, res, origin, cell = oracle.voxelize(lines, s, bounds=(lo32, hi32))
        ext = hi32 - lo32
        whole = np.ceil(ext / s32) * s32
        assert np.array_equal(res, (whole / s32).astype(np.int32)) and np.array_equal(origin, lo32) and cell == s32
    # a grid of 2^31 cells or more is refused
    with pytest.raises(ValueError):
        oracle.voxelize(np.zeros((1, 8), np.float32), 0.01, bounds=((0, 0, 0), (13.0, 13.0, 13.0)))


def test_axis_aligned_line_closed_form(oracle):
    s, r = 0.25, 0.0625
    a, b = np.array([0.375, 1.0, 1.5], np.float32), np.array([2.25, 1.0, 1.5], np.float32)
    lines = np.zeros((1, 8), np.float32)
    lines[0, :3], lines[0, 4:7] = a, b
    bounds = (np.zeros(3, np.float32), np.full(3, 3.0, np.float32))
    bits, res, origin, _ = oracle.voxelize(lines, s, r, bounds=bounds)
    assert list(res) == [12, 12, 12]
    got = _bits_to_cells(bits, res)
    want = np.zeros_like(got)
    lo, hi = a - r, b + r
    i0, i1 = np.floor(lo / s).astype(int), np.floor(hi / s).astype(int)
    want[i0[2]:i1[2] + 1, i0[1]:i1[1] + 1, i0[0]:i1[0] + 1] = True
    assert np.array_equal(got, want)
    # reference sweep: x = floor(step start / s) of every step t = a.x - r, + s, ... < b.x + r
    bits, res, _, _ = oracle.voxelize(lines, s, r, bounds=bounds, reference_sweep=True)
    got = _bits_to_cells(bits, res)
    starts = np.arange(lo[0], hi[0], s)
    want = np.zeros_like(got)
    for t in starts:
        want[i0[2]:i1[2] + 1, i0[1]:i1[1] + 1, int(np.floor(t / s))] = True
    assert np.array_equal(got, want)
    assert not got[:, :, i1[0]].any()  # the last partial cell along x is left out


def _segment_distance(p, a, b):
    ab = b - a
    t = np.clip(((p - a) @ ab) / (ab @ ab), 0.0, 1.0)
    return np.linalg.norm(p - (a + t[:, None] * ab[None, :]), axis=1)


@pytest.mark.parametrize("plane", [0, 1, 2])
def test_capsules_in_a_coordinate_plane_are_covered(oracle, plane):
    rng = np.random.default_rng(7 + plane)
    s, r = 0.1, 0.03
    n = 40
    a = rng.uniform(0.5, 3.5, (n, 3)).astype(np.float32)
    b = (a + rng.normal(size=(n, 3)) * 0.6).astype(np.float32)
    b[:, plane] = a[:, plane]  # each line lies in a plane normal to axis `plane`
    lines = np.zeros((n, 8), np.float32)
    lines[:, :3], lines[:, 4:7] = a, b
    bounds = (np.zeros(3, np.float32), np.full(3, 4.0, np.float32))
    bits, res, origin, cell = oracle.voxelize(lines, s, r, bounds=bounds)
    grid = _bits_to_cells(bits, res)
    for i in range(n):
        # points of the capsule: segment points plus offsets of length < r
        t = rng.random(400)[:, None]
        v = rng.normal(size=(400, 3))
        v *= (0.999 * r * rng.random(400) ** (1 / 3) / np.linalg.norm(v, axis=1))[:, None]
        p = a[i] + t * (b[i] - a[i]) + v
        c = np.floor((p - origin) / cell).astype(int)
        inside = np.all((c >= 0) & (c < res), axis=1)
        assert grid[c[inside, 2], c[inside, 1], c[inside, 0]].all(), f"line {i}"
    # nothing far from every capsule: a filled cell's centre is within r + one step box of a segment
    zz, yy, xx = np.nonzero(grid)
    centres = (np.stack([xx, yy, zz], 1) + 0.5) * cell + origin
    dist = np.min(np.stack([_segment_distance(centres, a[i], b[i]) for i in range(n)]), axis=0)
    assert dist.max() < r / np.sqrt(0.5) + 2.0 * s * np.sqrt(3.0)


def test_reference_sweep_fills_one_column_for_y_major_lines(oracle):
    """A y-major line at constant x = 1.0 with radius 0.05: every sweep box spans x in
    [0.95, 1.05], cells 9 and 10; the reference's loop keeps x at the box minimum (cell 9)."""
    lines = np.array([[1.0, 0.2, 1.01, 0, 1.0, 3.1, 1.33, 0]], np.float32)
    bounds = (np.zeros(3, np.float32), np.full(3, 4.0, np.float32))
    full = _bits_to_cells(*oracle.voxelize(lines, 0.1, 0.05, bounds=bounds)[:2])
    ref = _bits_to_cells(*oracle.voxelize(lines, 0.1, 0.05, bounds=bounds, reference_sweep=True)[:2])
    assert list(np.unique(np.nonzero(ref)[2])) == [9] and list(np.unique(np.nonzero(full)[2])) == [9, 10]
    assert np.array_equal(ref, full & (np.arange(full.shape[2]) == 9)[None, None, :])


def test_levels_of_detail_are_or_reductions(oracle):
    rng = np.random.default_rng(3)
    lines = np.zeros((60, 8), np.float32)
    lines[:, :3] = rng.uniform(-3, 3, (60, 3))
    lines[:, 4:7] = lines[:, :3] + rng.normal(size=(60, 3)) * 0.7
    bits0, res0, org0, cell0 = oracle.voxelize(lines, 0.1, 0.04)
    g = _bits_to_cells(bits0, res0)
    for lod in (1, 2, 3):
        bits, res, org, cell = oracle.voxelize(lines, 0.1, 0.04, lod=lod)
        k = 2 ** lod
        assert cell == np.float32(0.1) * k and np.array_equal(org, org0)
        assert list(res) == [(int(x) + k - 1) // k for x in res0]
        pad = np.zeros(tuple(int(x) * k for x in res[::-1]), bool)
        pad[:g.shape[0], :g.shape[1], :g.shape[2]] = g
        want = pad.reshape(res[2], k, res[1], k, res[0], k).any(axis=(1, 3, 5))
        assert np.array_equal(_bits_to_cells(bits, res), want)


def _random_grid(rng, res, density):
    n = int(np.prod(res))
    cells = rng.random(n) < density
    bits = np.packbits(cells, bitorder="little")
    bits = np.concatenate([bits, np.zeros((-len(bits)) % 4, np.uint8)]).view(np.uint32)
    return bits, cells.reshape(res[2], res[1], res[0])


def test_walk_matches_brute_force(oracle):
    rng = np.random.default_rng(11)
    res = np.array([23, 17, 29], np.int32)
    origin, cell = np.array([-1.5, 0.25, 2.0], np.float32), np.float32(0.37)
    bits, _ = _random_grid(rng, res, 0.04)
    S = oracle.Scene()
    S.add_voxels(bits, res, origin, cell)
    n = 20000
    centre = origin + res * cell / 2
    o = centre + rng.normal(size=(n, 3)) * 8.0
    tgt = origin + rng.random((n, 3)) * res * cell
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 0.001), d, np.full((n, 1), 1e4)], 1).astype(np.float32)
    inside = rng.random(n) < 0.2  # some rays start inside the grid
    rays[inside, :3] = tgt[inside] - d[inside] * 0.05
    walk = S.trace(rays)
    brute = S.trace(rays, flags=oracle.F_BRUTE)
    hit_w, hit_b = walk["prim"] != 0xFFFFFFFF, brute["prim"] != 0xFFFFFFFF
    assert hit_w.sum() > n // 3
    agree = (walk["prim"] == brute["prim"]) & (hit_w == hit_b)
    print(f"walk vs brute force: {agree.mean():.6f} agree")
    assert agree.mean() > 0.9995
    both = hit_w & hit_b & agree
    assert np.abs(walk["t"][both] - brute["t"][both]).max() < 1e-4
    # unit face normals against the ray
    nrm = walk["n"][hit_w]
    assert np.all(np.abs(nrm).sum(1) == 1.0) and np.all((nrm * rays[hit_w, 4:7]).sum(1) <= 0.0)
    # any-hit agrees on occlusion
    anyh = S.trace(rays, flags=oracle.F_ANY_HIT)
    assert np.array_equal(anyh["prim"] != 0xFFFFFFFF, hit_w)


def test_walk_closed_forms_along_axes(oracle):
    res = np.array([8, 4, 2], np.int32)
    origin, cell = np.array([1.0, 2.0, 3.0], np.float32), np.float32(0.5)
    grid = np.zeros((2, 4, 8), bool)
    grid[1, 2, 5] = True
    grid[0, 3, 1] = True
    bits = np.packbits(grid.reshape(-1), bitorder="little")
    bits = np.concatenate([bits, np.zeros((-len(bits)) % 4, np.uint8)]).view(np.uint32)
    S = oracle.Scene()
    S.add_voxels(bits, res, origin, cell)
    rays = np.array([
        [0.0, 3.25, 3.75, 0.001, 1, 0, 0, 100],    # +x through row (y 2, z 1): cell x 5 at x = 3.5
        [9.0, 3.25, 3.75, 0.001, -1, 0, 0, 100],   # -x: enters cell x 5 from x = 4.0
        [1.75, 0.0, 3.25, 0.001, 0, 1, 0, 100],    # +y through column (x 1, z 0): cell y 3 at y = 3.5
        [2.25, 0.0, 3.25, 0.001, 0, 1, 0, 100],    # +y through an empty column
        [3.75, 3.25, 3.75, 0.001, 0, 0, 1, 100],   # starts inside the filled cell (5, 2, 1)
    ], np.float32)
    h = S.trace(rays)
    assert list(h["prim"][:3]) == [5 + 2 * 8 + 32, 5 + 2 * 8 + 32, 1 + 3 * 8]
    assert list(h["t"][:3]) == [3.5, 5.0, 3.5]
    assert h["prim"][3] == 0xFFFFFFFF
    assert h["prim"][4] == 5 + 2 * 8 + 32 and h["t"][4] == np.float32(0.001)
    assert np.array_equal(h["n"][0], [-1, 0, 0]) and np.array_equal(h["n"][1], [1, 0, 0])
    assert np.array_equal(h["n"][2], [0, -1, 0]) and np.array_equal(h["n"][4], [0, 0, -1])
    # a translated instance: the same hits with the ray moved along
    S2 = oracle.Scene()
    S2.add_voxels(bits, res, origin, cell, transform=np.array([[1, 0, 0, 2.0], [0, 1, 0, -1.0], [0, 0, 1, 0.5]]))
    moved = rays.copy()
    moved[:, :3] += np.array([2.0, -1.0, 0.5], np.float32)
    h2 = S2.trace(moved)
    assert np.array_equal(h2["prim"], h["prim"]) and np.array_equal(h2["t"], h["t"])


def test_brick_jump_equals_cell_walk():
    """The device walk crosses an empty 16x16x16 block or 4x4x4 brick in one move
    (trace_impl.hpp, voxel_walk). tests/cpp/voxel_brick_walk.cpp restates those moves and the
    cell-by-cell DDA in host C++ with the same float arithmetic, and compares hit, cell, entry t
    (bits) and face over 2.4M rays on 60 random grids; with the tie rule flipped on purpose
    ("flip") it must find mismatches."""
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(here, "cpp", "voxel_brick_walk")
    subprocess.run(["make", "-s", "-C", os.path.join(here, "cpp"), exe], check=True)

    def mismatches(*args):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300, check=True)
        last = r.stdout.strip().splitlines()[-1].split()
        assert last[0] == "2400000" and int(last[2]) > 100000, r.stdout[-500:]
        return int(last[4])

    assert mismatches() == 0
    assert mismatches("flip") > 0

