"""Set-up checks of implicit diffusion (the reference's _IMPDIFF build) on a
host-only context: the cases the tensor-product and the MG-PCG Helmholtz
solves take, and the clear errors for the ones they do not."""
import pytest

import cases
import impdiff_cases


def host_ctx(tmp_path, blocks, fft_axis, **dns):
    import snac_amd
    d, b = cases.write_case(str(tmp_path), blocks, 100.0, **dns)
    return snac_amd.Context(d, b, fft_axis, device=-1)


@pytest.mark.parametrize("geom,fft_axis", [("channel", 3), ("channel", 1), ("cavity", 0), ("cavity", 3),
                                           ("duct", 3), ("box", 1), ("box", 0)])
def test_accepted(tmp_path, geom, fft_axis):
    blocks, dns = impdiff_cases.GEOMS[geom]()
    ctx = host_ctx(tmp_path, blocks, fft_axis, **dns)
    ctx.set_impdiff(True)
    ctx.set_impdiff(False)
    ctx.close()


@pytest.mark.parametrize("geom", ["lchannel", "bfs"])
def test_multigrid_cases_accepted(tmp_path, geom):
    """unions of blocks that are not boxes: every field's Helmholtz operator
    assembles on the MG-PCG path (on the step, v's link from the lower
    block's top face to the inlet's wall face drops out)"""
    blocks = cases.lchannel_blocks(4, 8) if geom == "lchannel" else cases.bfs_blocks((6, 4), (8, 4, 6), 4)
    ctx = host_ctx(tmp_path, blocks, 3)
    ctx.set_impdiff(True)
    ctx.close()


def test_multigrid_case_without_fft_axis_rejected(tmp_path):
    import snac_amd
    ctx = host_ctx(tmp_path, cases.lchannel_blocks(4, 8), 0)
    with pytest.raises(snac_amd.SnacError, match="needs an FFT axis"):
        ctx.set_impdiff(True)
    ctx.close()


def test_staggered_neumann_transform_rejected(tmp_path):
    """u with Neumann walls in x on a uniform x axis that the plan transforms:
    the staggered NN pair has no matching transform"""
    import snac_amd
    F1 = ("F", 1)
    blk = cases._block((8, 8, 8), (0, 0, 0), (1, 1, 1), vel=[["N", "N"], ["D", "D"], [F1, F1]],
                       pre=[["N", "N"], ["N", "N"], [F1, F1]], gr=(0.0, 1.2, 0.0))
    ctx = host_ctx(tmp_path, [blk], 3)
    with pytest.raises(snac_amd.SnacError, match="staggered"):
        ctx.set_impdiff(True)
    ctx.close()


def test_scalars_accepted(tmp_path):
    blocks, _ = impdiff_cases.channel((8, 8, 8))
    blocks[0]["cbcscal"] = [[["F", "F"], ["D", "D"], ["F", "F"]]]
    blocks[0]["bcscal"] = [[[1.0, 1.0], [0.0, 1.0], [1.0, 1.0]]]
    ctx = host_ctx(tmp_path, blocks, 3, nscal=1,
                   scalar=dict(alphai=[100.0], beta=0.0, iniscal=["zer"], ssource=[0.0], is_sforced=[False],
                               scalf=[0.0]))
    ctx.set_impdiff(True)
    ctx.close()
