"""Text reals parsed to float32 as the reference's list-directed reads do.

A Fortran ``read(..., *) x`` into a default real rounds the decimal once, to
the nearest float32.  Parsing to float64 and then rounding to float32 differs
where the float64 value lies exactly halfway between two float32 values;
adjoint-source files written from float32 seismograms times 1e6 hold such
values (155 of the 90,000 samples of the 48x48x24 gradient test's files).
Expected bit patterns: what amdflang's runtime reads for each token."""
import numpy as np

from ses3d_amd import formats as F

CASES = [
    ("1.0000000596046448", 0x3F800001), ("-1.0000000596046448", 0xBF800001),
    ("1.0000000596046447", 0x3F800000), ("0.1", 0x3DCCCCCD), ("1e-45", 0x00000001),
    ("7.006492321624085e-46", 0x00000000), ("-8.019313686413083e-21", 0x9E177B05),
    ("-2.4836159486164136e-21", 0x9D3BA81F), ("4.204452222907811e-21", 0x1D9ED6F7),
    ("1.1360465009370254e-20", 0x1E5697D7), ("-7.847802956279041e-21", 0x9E143DA5),
    ("1.6000190939323727e-20", 0x1E971E17), ("2.5d0", 0x40200000),
]


def test_reals_round_once_to_float32():
    toks = [t for t, _ in CASES]
    want = np.array([h for _, h in CASES], np.uint32)
    np.testing.assert_array_equal(F.f32_array(toks).view(np.uint32), want)
    for t, h in CASES:
        assert int(F.f32(t).view(np.uint32)) == h, t
    # rounding through float64 gets some of them wrong
    naive = np.array([np.float32(float(t.replace("d", "e"))) for t in toks]).view(np.uint32)
    assert int((naive != want).sum()) == 8


def test_adjoint_source_file_values(tmp_path):
    """An ad_src file holding halfway values reads back as the reference reads it."""
    rows = np.array([[1.0000000596046448, -8.019313686413083e-21, 0.5]] * 2)
    p = tmp_path / "ad_src_1"
    F.write_ad_src(str(p), (50.0, 20.0, 0.0), rows[::-1].T)
    got = F.read_ad_src(str(p), 2).view(np.uint32)
    np.testing.assert_array_equal(got[:, 0], [0x3F800001] * 2)
    np.testing.assert_array_equal(got[:, 1], [0x9E177B05] * 2)
