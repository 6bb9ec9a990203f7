"""Default tile of SRParams: 128-cell tiles where the box has at least 512 of
them (two fused-kernel workgroups per CU), 64-cell tiles otherwise."""
import pytest


@pytest.mark.parametrize("n,tile", [((256, 256), (16, 8)), ((128, 128), (8, 8)),
                                    ((24, 32), (8, 8)), ((512, 256, 256), (8, 4, 4)),
                                    ((128, 64, 64), (8, 4, 4)), ((16, 12, 8), (4, 4, 4))])
def test_default_tile(n, tile):
    from entity_amd.engine import SRParams
    assert SRParams(dim=len(n), n=n).tile == tile
    assert SRParams(dim=len(n), n=n, tile=(4,) * len(n)).tile == (4,) * len(n)
