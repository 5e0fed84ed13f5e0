"""The oracle's screened-grid restatement (Morton order, tile boxes, neighbor
list; octree.hpp:70-389) against what geometry fixes: the interleaving read
back bit by bit, the octant order of a 2 x 2 x 2 lattice, boxes that hold
their points, and spheres that touch, miss or contain a box."""
import numpy as np

from oracle import oracle as O


def _deinterleave(code, axis):
    """Bits 3 b + (2 - axis) of code, gathered into an integer."""
    out = np.zeros_like(code)
    for b in range(21):
        out |= ((code >> np.uint64(3 * b + 2 - axis)) & np.uint64(1)) << np.uint64(b)
    return out


def test_morton_codes_read_back():
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(500, 3)) * [3.0, 1.0, 0.2]
    code = O.morton_codes(pts)
    lo, hi = pts.min(0), pts.max(0)
    for d in range(3):
        q = np.floor((pts[:, d] - lo[d]) / (hi[d] - lo[d]) * 2097151.0).astype(np.uint64)
        assert np.array_equal(_deinterleave(code, d), q)
    assert code.max() < (1 << 63)


def test_morton_octant_order():
    """Corners of a cube: z varies fastest, then y, then x."""
    corners = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
    rng = np.random.default_rng(1)
    shuffled = corners[rng.permutation(8)]
    perm, boxes = O.bounding_boxes(shuffled, 1)
    assert np.array_equal(shuffled[perm], corners)
    assert np.array_equal(boxes[:, :3], corners) and np.array_equal(boxes[:, 3:], corners)


def test_bounding_boxes_hold_their_points_ragged():
    rng = np.random.default_rng(2)
    pts = rng.uniform(-4, 4, size=(1000, 3))
    per = 64
    perm, boxes = O.bounding_boxes(pts, per)
    assert np.array_equal(np.sort(perm), np.arange(1000))
    sp = pts[perm]
    assert boxes.shape == (16, 6)  # 15 full boxes and one of 40 points
    for b in range(16):
        blk = sp[b * per:(b + 1) * per]
        assert np.array_equal(boxes[b, :3], blk.min(0))
        assert np.array_equal(boxes[b, 3:], blk.max(0))
    # Morton order keeps consecutive points close: the boxes are far smaller
    # than the cloud
    assert np.median(np.prod(boxes[:, 3:] - boxes[:, :3], axis=1)) < 0.2 * 8.0 ** 3


def test_neighbor_list_touching_missing_containing():
    box = np.array([[0.0, 0.0, 0.0, 1.0, 1.0, 1.0]])
    c = np.array([[0.5, 0.5, 0.5],   # inside, tiny sphere
                  [2.0, 0.5, 0.5],   # touches the x = 1 face (distance exactly 1)
                  [2.0, 0.5, 0.5],   # just short of it
                  [2.0, 2.0, 2.0],   # touches the corner (distance sqrt 3)
                  [-3.0, 0.5, 0.5]])  # far
    r = np.array([1e-3, 1.0, 0.999999, np.sqrt(3.0), 1.0])
    off, nbr = O.neighbor_list(box, c, r)
    assert off.tolist() == [0, 3] and nbr.tolist() == [0, 1, 3]


def test_neighbor_list_matches_all_pairs_distances():
    rng = np.random.default_rng(3)
    pts = rng.uniform(-5, 5, size=(700, 3))
    perm, boxes = O.bounding_boxes(pts, 50)
    c = rng.uniform(-6, 6, size=(40, 3))
    r = rng.uniform(0.5, 3.0, size=40)
    off, nbr = O.neighbor_list(boxes, c, r)
    sp = pts[perm]
    for b in range(len(boxes)):
        got = nbr[off[b]:off[b + 1]]
        assert np.all(np.diff(got) > 0)
        # a sphere meets the box iff its centre is within r of the box: a
        # sphere holding one of the box's points always does
        blk = sp[b * 50:(b + 1) * 50]
        near = np.nonzero((np.linalg.norm(blk[:, None, :] - c[None], axis=2) <= r).any(0))[0]
        assert set(near.tolist()) <= set(got.tolist())
        # and one that meets it comes within r of the box's nearest point
        q = np.clip(c[got], boxes[b, :3], boxes[b, 3:])
        assert np.all(np.linalg.norm(q - c[got], axis=1) <= r[got] * (1 + 1e-15))
