"""Hand-derived known answers for single rays of the Siddon pair.

Each case is one ray (view angle 0, one detector cell) through a small volume
of 1 mm voxels, placed so that the slab walk meets round numbers.  The
expected weights were worked out by hand from the reference's slab rule
(nefct/cc/siddon_3d/siddon.cu.cc:41-176 for x slabs, :186-316 for y slabs,
the adjoint's copies at :480-764) and from the exact line lengths; they are
literals here, not recomputed by any formula of the code under test.

Geometry (siddon-op names): source at (-SID, 0, 0); the ray's detector cell is
at (SAD - SID, a, b) with na = nb = 1, da = db = 1, so ai = a + 1/2 and
bi = b + 1/2.  All cases use SID = 100 and SAD = 200 unless stated.

Weights are mm of ray per voxel (adjoint with a unit sinogram; the
reference's forward divides them by L^2, L in voxel units, Q1).  Volumes are
indexed [z, y, x].
"""
import math

SID, SAD = 100.0, 200.0

# Q2, the mixed-sign double crossing, in each branch of each walk.
#
# Case A, x slabs, ry <= rz: the slab x in [-1/2, 1/2] has y running 1.75 ->
# 2.25 (ky = +1/2) and z running 2.125 -> 1.625 (kz = -1/2).  The reference
# sorts each axis on its own (yy1 = 1.75, zz1 = 1.625), so ry = 0.5 and
# rz = 0.75 are measured from opposite faces, and it credits (y1,z1) ry,
# (y2,z1) rz - ry, (y2,z2) 1 - rz.  Exactly, y crosses 2 at 1/2 of the slab
# and z crosses 2 at 1/4: (y1,z2) 1/4, (y1,z1) 1/4, (y2,z1) 1/2.
# Case B, x slabs, ry > rz: y 1.625 -> 2.125 (ry = 0.75), z 2.25 -> 1.75
# (zz1 = 1.75, rz = 0.5): reference (y1,z1) rz, (y1,z2) ry - rz, (y2,z2)
# 1 - ry; exact: z crosses at 1/2, y at 3/4.
# Cases C and D repeat B and A on a y slab (|yd| = 400 > |xd| = 200), with x
# in place of y.
W15 = math.sqrt(1.5)   # sqrt(1 + 1/4 + 1/4) mm per slab

Q2_CASES = {
    "x_walk_ry_le_rz": dict(
        grid=(1, 4, 4), center=(0.0, 50.0, -49.875), a=100.0, b=-100.0,
        ref={(1, 1, 0): 0.5 * W15, (1, 2, 0): 0.25 * W15, (2, 2, 0): 0.25 * W15},
        exact={(2, 1, 0): 0.25 * W15, (1, 1, 0): 0.25 * W15, (1, 2, 0): 0.5 * W15}),
    "x_walk_ry_gt_rz": dict(
        grid=(1, 4, 4), center=(0.0, 50.125, -50.0), a=100.0, b=-100.0,
        ref={(1, 1, 0): 0.5 * W15, (2, 1, 0): 0.25 * W15, (2, 2, 0): 0.25 * W15},
        exact={(2, 1, 0): 0.5 * W15, (1, 1, 0): 0.25 * W15, (1, 2, 0): 0.25 * W15}),
    "y_walk_rx_gt_rz": dict(
        grid=(4, 1, 4), center=(0.125, 200.0, -100.0), a=400.0, b=-200.0,
        ref={(1, 0, 1): 0.5 * W15, (2, 0, 1): 0.25 * W15, (2, 0, 2): 0.25 * W15},
        exact={(2, 0, 1): 0.5 * W15, (1, 0, 1): 0.25 * W15, (1, 0, 2): 0.25 * W15}),
    "y_walk_rx_le_rz": dict(
        grid=(4, 1, 4), center=(0.0, 200.0, -99.875), a=400.0, b=-200.0,
        ref={(1, 0, 1): 0.5 * W15, (1, 0, 2): 0.25 * W15, (2, 0, 2): 0.25 * W15},
        exact={(2, 0, 1): 0.25 * W15, (1, 0, 1): 0.25 * W15, (1, 0, 2): 0.5 * W15}),
}
# the reference's forward scale 1/L^2 for these rays (L = |D - S| in voxels)
Q2_L2 = {"x_walk_ry_le_rz": 200.0 ** 2 + 100.0 ** 2 + 100.0 ** 2,
         "x_walk_ry_gt_rz": 200.0 ** 2 + 100.0 ** 2 + 100.0 ** 2,
         "y_walk_rx_gt_rz": 200.0 ** 2 + 400.0 ** 2 + 200.0 ** 2,
         "y_walk_rx_le_rz": 200.0 ** 2 + 400.0 ** 2 + 200.0 ** 2}

# Q3 (siddon.cu.cc:472): the adjoint skips rays whose in-plane length is below
# 10 voxels.  A ray along the x axis (a = b = 0) from x = -4 through a
# 4 x 2 x 2 volume centred at (0, 1/2, 1/2): each x slab credits voxel
# (0, 0, ix) with 1 mm.  Its in-plane length is SAD.
Q3_CASE = dict(grid=(4, 2, 2), center=(0.0, 0.5, 0.5), SID=4.0, a=0.0, b=0.0,
               hit={(0, 0, 0): 1.0, (0, 0, 1): 1.0, (0, 0, 2): 1.0, (0, 0, 3): 1.0})
Q3_SAD_SKIPPED, Q3_SAD_KEPT = 9.9, 10.1

# Q6 (:41 '>=' in the forward, :480 '>' in the adjoint): a ray with
# |xd| = |yd| = 200 exactly (a = SAD), kz = -1/2, through a 1 x 1 x 4 column
# centred at (0, 100.375, -50).  The forward walks x slabs: y runs -0.375 ->
# 0.625 (ry = 0.375), z 2.25 -> 1.75 (rz = 0.5); the reference credits
# (cy2=0, cz1=1) rz - ry = 1/8 and (0, 2) 1 - rz = 1/2 ((cy1 = -1) is out).
# The adjoint with the '>' rule walks the y slab instead: x runs 0.375 ->
# 1.375 (rx = 0.625), z 2.0625 -> 1.5625 (rz = 0.875); it credits (cx1 = 0,
# cz1 = 1) rx = 5/8 and nothing else (cx2 = 1 is out).  Exactly the column
# holds z = 1 for 1/2 and z = 2 for 1/8 of a slab.  sqrt(1 + 1 + 1/4) = 3/2.
Q6_CASE = dict(grid=(1, 1, 4), center=(0.0, 100.375, -50.0), a=200.0, b=-100.0,
               fwd_x_walk={(1, 0, 0): 0.125 * 1.5, (2, 0, 0): 0.5 * 1.5},
               adj_y_walk={(1, 0, 0): 0.625 * 1.5},
               exact={(1, 0, 0): 0.5 * 1.5, (2, 0, 0): 0.125 * 1.5})
Q6_L2 = 200.0 ** 2 + 200.0 ** 2 + 100.0 ** 2   # L = 300


def det_args(a, b):
    """(na, da, ai, nb, db, bi) of the single detector cell at (a, b)."""
    return 1, 1.0, a + 0.5, 1, 1.0, b + 0.5


def dense(grid, weights):
    """[nz, ny, nx] array holding the listed weights (numpy float64)."""
    import numpy as np
    nx, ny, nz = grid
    out = np.zeros((nz, ny, nx))
    for (z, y, x), w in weights.items():
        out[z, y, x] = w
    return out
