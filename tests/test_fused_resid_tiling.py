"""Host-side model of how the first-half residuals of a 3D level are split
between the pre-smoothing pair sweep (FR variant of k_rbsweep3p) and
k_resid_edges3: every (plane, row) of unknowns must be covered exactly once.

The model restates the index rules of both kernels in jsolver_amd/csrc/
jsv_kernels.hpp: a tile of ty rows starting at y0 = 2 + t*ty and a z-chunk
[z0, z1) with z1 = min(z0 + zchunk, mz + 1) do rows y0+1 .. y0+ty-2 (up to my)
and planes z0+1 .. z1-2 in the sweep; the edge pass's blocks decode as
(row pair at a tile boundary, plane) then (plane pair at a chunk boundary,
row), skipping tile edge rows in the second part.  The GPU parity tests
check the kernels themselves against the oracle; this checks the
decomposition over many more tilings than those tests can run.

Marching direction (dn_in of k_rbsweep3p): bit 0 marches every chunk
downward in the mirrored plane coordinate u (stored plane mz + 2 - u), the
chunks laid out in u; bit 1 alternates the direction per chunk, whose bounds
stay those of the upward march in stored planes.  The host hands the edge
pass the mirror (its dn) only in the first case.
"""
import itertools


def fused_points(my, mz, ty, zchunk, dn_in=0):
    ytiles = (my - 1 + ty - 1) // ty
    zchunks = (mz - 1 + zchunk - 1) // zchunk
    out = set()
    for t in range(ytiles):
        y0 = 2 + t * ty
        for c in range(zchunks):
            alt = (dn_in & 2) != 0
            dn = (dn_in & 1) ^ ((c & 1) if alt else 0)
            z0 = 2 + c * zchunk
            z1 = min(z0 + zchunk, mz + 1)
            if alt and dn:
                z0, z1 = mz + 3 - z1, mz + 3 - z0
            for r in range(2, ty):              # thread row r: y = y0 - 1 + r
                y = y0 - 1 + r
                if y > my:
                    continue
                for u in range(z0 + 1, z1 - 1):
                    out.add((mz + 2 - u if dn else u, y))
    return out


def edge_points(my, mz, ty, zchunk, dn=0):
    ytiles = (my - 1 + ty - 1) // ty
    zchunks = (mz - 1 + zchunk - 1) // zchunk
    n_a = (ytiles + 1) * (mz - 1)
    n_b = (zchunks + 1) * (my - 1)
    out = []
    for b in range(n_a + n_b):
        for h in (0, 1):
            if b < n_a:
                j = b // (mz - 1)
                z = 2 + (b - j * (mz - 1))
                if h == 0:
                    y = 2 + j * ty - 1
                    if j == 0 or y > my:
                        continue
                else:
                    y = 2 + j * ty
                    if j == ytiles:
                        continue
            else:
                bb = b - n_a
                j = bb // (my - 1)
                y = 2 + (bb - j * (my - 1))
                ry = (y - 2) % ty
                if ry == 0 or ry == ty - 1:
                    continue
                if h == 0:
                    if j == 0:
                        continue
                    z = min(2 + j * zchunk - 1, mz)
                    if z == 2 + (j - 1) * zchunk:
                        continue
                else:
                    z = 2 + j * zchunk
                    if j == zchunks:
                        continue
            out.append((mz + 2 - z if dn else z, y))
    return out


def test_sweep_and_edge_pass_cover_every_point_once():
    for my, mz, ty, zchunk in itertools.product([3, 4, 7, 16, 31, 64, 255], [3, 7, 16, 31, 255],
                                                [4, 5, 6, 10, 14], [1, 2, 3, 5, 20, 29, 51, 255]):
        if zchunk > mz:
            continue
        everything = {(z, y) for z in range(2, mz + 1) for y in range(2, my + 1)}
        for dn_in in (0, 1, 2, 3):
            key = (my, mz, ty, zchunk, dn_in)
            fused = fused_points(my, mz, ty, zchunk, dn_in)
            edge = edge_points(my, mz, ty, zchunk, dn_in & 1 if dn_in < 2 else 0)
            assert all(2 <= z <= mz and 2 <= y <= my for z, y in edge), key
            assert len(edge) == len(set(edge)), ("edge point twice", key)
            assert not fused & set(edge), ("point in both", key)
            assert fused | set(edge) == everything, ("point missed", key)


def test_alternating_chunks_store_the_upward_march_planes():
    """With alternating chunk directions each workgroup stores exactly the
    planes its chunk holds in the upward march, so the edge pass needs no
    mirror; a variant that forgot to remap the downward chunks' bounds
    would not (it would store the mirrored chunk's planes)."""
    for mz, zchunk in itertools.product([7, 16, 31, 255], [2, 3, 5, 29]):
        if zchunk > mz:
            continue
        for dn_in in (2, 3):
            assert fused_points(9, mz, 5, zchunk, dn_in) == fused_points(9, mz, 5, zchunk, 0), (mz, zchunk)
    # mirroring a downward chunk without remapping its bounds would store
    # other planes (chunk 1 of 29 planes at mz = 256)
    z0, z1 = 2 + 29, 2 + 2 * 29
    assert {256 + 2 - u for u in range(z0 + 1, z1 - 1)} != set(range(z0 + 1, z1 - 1))
