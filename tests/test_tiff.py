"""16-bit TIFF stack reader, pinned against an independent TIFF codec
(Pillow) in both directions, plus ImageJ's contiguous large-stack layout."""
import numpy as np
import pytest

from imaging_fcs_1_52_amd.tiff import read_tiff_stack, write_tiff_stack

PIL = pytest.importorskip("PIL.Image")


def _stack(T=5, H=7, W=9, seed=0):
    return np.random.default_rng(seed).integers(0, 65536, (T, H, W), dtype=np.uint16)


def test_reads_pillow_multipage_stack(tmp_path):
    st = _stack()
    frames = [PIL.fromarray(f) for f in st]
    p = tmp_path / "pil.tif"
    frames[0].save(p, save_all=True, append_images=frames[1:])
    got = read_tiff_stack(str(p))
    assert got.shape == st.shape and np.array_equal(np.asarray(got), st)


@pytest.mark.parametrize("imagej", [False, True])
def test_written_stack_round_trips(tmp_path, imagej):
    st = _stack(T=6, H=5, W=8, seed=1)
    p = tmp_path / ("ij.tif" if imagej else "pages.tif")
    write_tiff_stack(str(p), st, imagej=imagej)
    got = read_tiff_stack(str(p))
    assert np.array_equal(np.asarray(got), st)
    # an independent reader sees the same first frame (and every page of
    # the one-page-per-frame layout)
    im = PIL.open(str(p))
    assert np.array_equal(np.asarray(im), st[0])
    if not imagej:
        for t in range(st.shape[0]):
            im.seek(t)
            assert np.array_equal(np.asarray(im), st[t])


def test_contiguous_stack_is_memory_mapped(tmp_path):
    st = _stack(T=4, H=3, W=4, seed=2)
    p = tmp_path / "mm.tif"
    write_tiff_stack(str(p), st, imagej=True)
    got = read_tiff_stack(str(p))
    assert isinstance(got, np.memmap)


def test_rejects_non_16_bit(tmp_path):
    p = tmp_path / "u8.tif"
    PIL.fromarray(np.zeros((4, 4), np.uint8)).save(p)
    with pytest.raises(ValueError):
        read_tiff_stack(str(p))
