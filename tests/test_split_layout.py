"""CPU check of the index arithmetic of the split plane-kernel applies (the
compact tile-boundary work layout and the node enumeration of the gather
pass), restated in numpy from aspect_amd/csrc/tile_common.h
(tile_bnd_index, xcd_order) and kernels_gather.hip (bnd_node, bnd_grid):
every velocity node of a level is finished exactly once, either inside its
tile (tile-interior nodes) or by one gather thread (tile-boundary nodes); the
compact index is a bijection of the 386 boundary nodes of a tile; every
gather thread's work offsets stay inside the array; the XCD block order is a
bijection."""
import numpy as np
import pytest


def tile_bnd_index(lx, ly, lz):
    if lz in (0, 8):
        return (81 if lz == 8 else 0) + 9 * ly + lx
    if ly in (0, 8):
        return 162 + (63 if ly == 8 else 0) + 9 * (lz - 1) + lx
    return 288 + (49 if lx == 8 else 0) + 7 * (lz - 1) + (ly - 1)


def xcd_order(b, nb):
    q, r, x, k = nb >> 3, nb & 7, b & 7, b >> 3
    return x * q + min(x, r) + k


def bnd_grid(nv):
    nzp = (nv[2] - 1) // 8 + 1
    full = (-(-nv[0] * nv[1] // 256), 3 * nzp)
    nyr, nxc = (nv[1] - 1) // 8 + 1, (nv[0] - 1) // 8 + 1
    n = nyr * nv[0] + nxc * (nv[1] - nyr)
    return full, (-(-n // 256), 3 * (nv[2] - nzp))


def bnd_nodes(nv, sparse):
    """(comp, gx, gy, gz) of every thread of the launch that passes the bound check."""
    gx_, gy_ = bnd_grid(nv)[sparse]
    nzp = (nv[2] - 1) // 8 + 1
    planes = nv[2] - nzp if sparse else nzp
    nyr, nxc = (nv[1] - 1) // 8 + 1, (nv[0] - 1) // 8 + 1
    out = []
    nb = gx_ * gy_
    for lin in range(nb):
        m = xcd_order(lin, nb)
        by, bx = divmod(m, gx_)
        comp, k = divmod(by, planes)
        gz = 8 * (k // 7) + k % 7 + 1 if sparse else 8 * k
        for t in range(256):
            i = bx * 256 + t
            if not sparse:
                if i >= nv[0] * nv[1]:
                    continue
                gy, gx = divmod(i, nv[0])
            else:
                nrow, ncol = nyr * nv[0], nxc * (nv[1] - nyr)
                if i >= nrow + ncol:
                    continue
                if i < nrow:
                    r, gx = divmod(i, nv[0])
                    gy = 8 * r
                else:
                    j = i - nrow
                    c, mm = divmod(j, nv[1] - nyr)
                    gx, gy = 8 * c, 8 * (mm // 7) + mm % 7 + 1
            out.append((comp, gx, gy, gz))
    return out


def test_tile_bnd_index_bijection():
    idx = sorted(tile_bnd_index(x, y, z) for z in range(9) for y in range(9) for x in range(9)
                 if 0 in (x, y, z) or 8 in (x, y, z))
    assert idx == list(range(386))


@pytest.mark.parametrize("nb", list(range(1, 70)) + [32768, 12347])
def test_xcd_order_bijection(nb):
    assert sorted(xcd_order(b, nb) for b in range(nb)) == list(range(nb))


@pytest.mark.parametrize("cells", [(8, 8, 8), (12, 8, 4), (6, 10, 14), (4, 4, 4), (2, 2, 2), (20, 12, 36)])
def test_every_node_finished_once(cells):
    nv = [2 * c + 1 for c in cells]
    nt = [-(-c // 4) for c in cells]
    count = np.zeros((3, nv[2], nv[1], nv[0]), dtype=int)
    # tile-interior nodes: rows with 1 <= ly, lz <= 7, nodes 1 <= lx <= 7, inside the level
    for tz in range(nt[2]):
        for ty in range(nt[1]):
            for tx in range(nt[0]):
                for lz in range(1, 8):
                    for ly in range(1, 8):
                        gy, gz = 8 * ty + ly, 8 * tz + lz
                        if gy >= nv[1] or gz >= nv[2]:
                            continue
                        for lx in range(1, 8):
                            gx = 8 * tx + lx
                            if gx < nv[0]:
                                count[:, gz, gy, gx] += 1
    work_size = nt[0] * nt[1] * nt[2] * 3 * 386
    for sparse in (0, 1):
        for comp, gx, gy, gz in bnd_nodes(nv, sparse):
            assert 0 <= gx < nv[0] and 0 <= gy < nv[1] and 0 <= gz < nv[2]
            assert gx % 8 == 0 or gy % 8 == 0 or gz % 8 == 0
            count[comp, gz, gy, gx] += 1
            # the gather's candidate tiles and work offsets
            for cz in ((gz >> 3) - 1, 8, gz % 8 == 0 and gz > 0), (gz >> 3, gz & 7, (gz >> 3) < nt[2]):
                for cy in ((gy >> 3) - 1, 8, gy % 8 == 0 and gy > 0), (gy >> 3, gy & 7, (gy >> 3) < nt[1]):
                    for cx in ((gx >> 3) - 1, 8, gx % 8 == 0 and gx > 0), (gx >> 3, gx & 7, (gx >> 3) < nt[0]):
                        if not (cz[2] and cy[2] and cx[2]):
                            continue
                        tile = cx[0] + nt[0] * (cy[0] + nt[1] * cz[0])
                        wi = (tile * 3 + comp) * 386 + tile_bnd_index(cx[1], cy[1], cz[1])
                        assert 0 <= wi < work_size
    assert (count == 1).all()
