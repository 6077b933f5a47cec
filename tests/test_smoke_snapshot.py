"""The snapshot __graft_entry__.smoke() builds has patches on every level it
asks for, so the smoke run exercises the patch deposit and not only the base
grid (a change of the synthetic generator could otherwise leave it empty)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import __graft_entry__ as entry  # noqa: E402


def test_smoke_snapshot_has_patches_on_every_level():
    snap = entry.smoke_snapshot()
    g = snap.geom
    assert g.nl == 3
    assert all(int(g.npatch[l]) > 0 for l in range(1, 4)), g.npatch
    assert snap.n == 24 ** 3 and g.nx == 32
