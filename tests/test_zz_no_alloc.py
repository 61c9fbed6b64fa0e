"""lee_write_contour_noalloc, the entry point that allocates no device memory, against
lee_write_contour through the C ABI: the same file, byte for byte, on two calls in a row (the
second on a smaller grid, so the shared staging array still holds the first call's larger
contents), and a clear error when a zone exceeds the staging array."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _files(path):
    names = sorted(os.listdir(path))
    return names, [open(os.path.join(path, n), "rb").read() for n in names]


def _both(s, tmp_path, tag):
    a, b = tmp_path / f"{tag}_alloc", tmp_path / f"{tag}_noalloc"
    a.mkdir()
    b.mkdir()
    s.write_contour(str(a))
    s.write_contour_noalloc(str(b))
    na, fa = _files(str(a))
    nb, fb = _files(str(b))
    return na, fa, nb, fb


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_contour_matches_the_allocating_writer(lee, tmp_path, prec):
    big = lee.deck.box(90, 70, 40, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=6, src=(0.0, 1.5, 0.5),
                       gauss_width=0.3, time_domain=40, out_difft=10)
    small = lee.deck.box(30, 26, 22, 0.05, pml=4, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.6, 0.5), gauss_width=0.25,
                         time_domain=40, out_difft=10)
    for tag, d, n in (("big", big, 25), ("small", small, 15)):
        print(f"[gpu no-alloc] fp{prec} {tag} {d.shape}", flush=True)
        s = lee.Solver(d, precision=prec)
        s.frame_setup(0)
        s.step(n)
        na, fa, nb, fb = _both(s, tmp_path, f"{tag}{prec}")
        s.close()
        assert na == nb and len(na) == 1
        assert fa == fb, tag
        # the file holds the subsampled pressure: 4 columns, values that moved off zero
        rows = [r.split() for r in fa[0].decode().splitlines() if r[:1] in "-0123456789"]
        vals = np.array([float(r[3]) for r in rows])
        assert np.abs(vals).max() > 0


def test_noalloc_contour_refuses_a_zone_beyond_the_staging_array(lee, tmp_path):
    # 1200 x 1200 x 20 subsamples to 301 x 301 x 6 = 543606 values > 524288
    d = lee.deck.box(1200, 1200, 20, 0.05, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 30.0, 0.5), gauss_width=0.3,
                     time_domain=2, out_difft=1)
    print("[gpu no-alloc] oversized zone", flush=True)
    s = lee.Solver(d, precision=32)
    s.frame_setup(0)
    with pytest.raises(lee.LeeError) as e:
        s.write_contour_noalloc(str(tmp_path))
    assert "exceeds the staging array" in str(e.value)
    assert os.listdir(str(tmp_path)) == []
    s.close()
