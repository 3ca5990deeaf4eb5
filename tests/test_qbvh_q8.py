"""The device node format (lr_node_q8): child boxes quantised to 8 bits per
plane on a per-node power-of-two grid, rounded outward.

Pinned on the CPU, before any GPU run:
- every decoded box contains the FP32 box of lr_node_export (conservative),
  and is tight: each plane is within one grid step of the FP32 plane, and the
  grid step is the smallest power of two that spans the node;
- a planted inward rounding is caught by that check;
- the oracle's traversal over the decoded nodes (oracle.decode_q8, the GPU's
  box tests bit for bit) returns exactly the hits of the FP32 traversal, which
  the GPU tests compare with.
"""
import numpy as np
import pytest

import liar_amd as la
import oracle

EMPTY = 0xFFFFFFFF


def scene(kind, seed):
    rng = np.random.default_rng(seed)
    sph = None
    if kind == "soup":
        v, t = la.random_triangles(4000, seed=seed)
    elif kind == "spheres":
        v, t = la.random_triangles(1000, seed=seed)
        sph = np.concatenate([rng.random((300, 3)), rng.uniform(0.002, 0.05, (300, 1))], 1).astype(np.float32)
    elif kind == "far":  # large coordinates, small triangles
        v, t = la.random_triangles(3000, seed=seed)
        v = (v * 40.0 + np.float32(5000.0)).astype(np.float32)
    elif kind == "flat":  # zero extent along z
        v, t = la.random_triangles(2000, seed=seed)
        v = v.copy()
        v[:, 2] = 0.25
    elif kind == "one":
        v, t = la.random_triangles(1, seed=seed)
    else:
        raise ValueError(kind)
    return v, t, sph


def boxes(nodes):
    """(n, 6, 4) float32 planes and (n, 4) uint32 children of lr_node bytes."""
    n = nodes.shape[0]
    planes = nodes[:, :96].copy().view(np.float32).reshape(n, 6, 4)
    child = nodes[:, 96:112].copy().view(np.uint32).reshape(n, 4)
    return planes, child


def containment_errors(nodes, q8):
    """Number of (node, child, plane) where the decoded plane is inside the FP32 box."""
    fp, child = boxes(nodes)
    dq, dchild = boxes(oracle.decode_q8(q8))
    assert np.array_equal(child, dchild)
    valid = np.repeat((child != EMPTY)[:, None, :], 6, axis=1)
    lo = np.zeros(6, bool)
    lo[0::2] = True
    inward = np.where(lo[None, :, None], dq > fp, dq < fp)
    return int((inward & valid).sum())


@pytest.mark.parametrize("kind", ["soup", "spheres", "far", "flat", "one"])
def test_decoded_boxes_contain_fp32_boxes_and_are_tight(kind):
    v, t, s = scene(kind, 3)
    b = la.Qbvh(None, v, t, s)
    nodes, _ = b.export()
    q8 = b.export_q8()
    assert q8.shape == (nodes.shape[0], 64)
    assert containment_errors(nodes, q8) == 0
    # tightness: within one grid step (plus the rounding of the decode)
    fp, child = boxes(nodes)
    dq, _ = boxes(oracle.decode_q8(q8))
    hdr = q8[:, :24].copy().view(np.float32)
    step = np.repeat(hdr[:, 3:6], 2, axis=1)[:, :, None]  # (n, 6, 1)
    valid = np.repeat((child != EMPTY)[:, None, :], 6, axis=1)
    slack = np.abs(dq - fp) - step - np.spacing(np.abs(fp)) * 2
    assert (slack[valid] <= 0).all(), float(slack[valid].max())
    # the grid step is the smallest power of two whose 255 steps span the node
    org = hdr[:, 0:3]
    scale = hdr[:, 3:6]
    assert (np.frexp(scale)[0] == 0.5).all()  # powers of two
    top = np.where(child[:, None, :] != EMPTY, fp[:, 1::2, :], -np.inf).max(axis=2)  # (n, 3)
    half = scale / 2
    has = (child != EMPTY).any(axis=1)[:, None] & (half >= np.float32(2.0 ** -126))
    spans_with_half = (org + np.float32(255) * half) >= top
    assert not (spans_with_half & has).any()


def test_planted_inward_rounding_is_caught():
    v, t, _ = scene("soup", 4)
    b = la.Qbvh(None, v, t)
    nodes, _ = b.export()
    q8 = b.export_q8()
    bad = q8.copy()
    # root, child 0: raise qlo_x by one step (rounds the plane inward)
    assert bad[0, 24] < 255
    bad[0, 24] += 1
    assert containment_errors(nodes, bad) > 0
    bad = q8.copy()
    bad[0, 28] -= 1  # qhi_x of child 0 one step down
    assert containment_errors(nodes, bad) > 0


@pytest.mark.parametrize("kind", ["soup", "spheres", "far", "flat"])
def test_traversal_over_decoded_nodes_equals_fp32_traversal(kind):
    v, t, s = scene(kind, 5)
    b = la.Qbvh(None, v, t, s)
    nodes, prims = b.export()
    dec = oracle.decode_q8(b.export_q8())
    r = la.random_rays(16384, seed=11)
    if kind == "far":
        r = r.copy()
        r[0:3] = r[0:3] * 40.0 + 5000.0
    ref = oracle.qbvh_closest(nodes, prims, r)
    got = oracle.qbvh_closest(dec, prims, r)
    for a, c in zip(ref[:4], got[:4]):
        np.testing.assert_array_equal(a.view(np.uint32), c.view(np.uint32))
    assert (ref[0] != oracle.MISS).any()
    # looser boxes cost extra node visits, within reason
    assert ref[4] <= got[4] * 1.02 and got[4] < 1.6 * ref[4], (ref[4], got[4])
    r2 = la.random_rays(16384, seed=12, tfar_lo=0.01, tfar_hi=2.0)
    if kind == "far":
        r2 = r2.copy()
        r2[0:3] = r2[0:3] * 40.0 + 5000.0
        r2[6:8] *= 40.0
    np.testing.assert_array_equal(oracle.qbvh_any(nodes, prims, r2)[0], oracle.qbvh_any(dec, prims, r2)[0])
