"""MatrixMarket reader (host code of the library; runs without a GPU):
the entries of small files of every supported kind, against hand-written
expectations that follow the reference's loadMM
(examples/load_mm_and_compare.cpp:287-428)."""
import numpy as np
import pytest

from spc5_amd import mm
from spc5_amd._lib import Spc5Error


def write(tmp_path, text, name="a.mtx"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_general_real_with_comments(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real general\n% c1\n%c2\n3 4 3\n1 1 2.5\n3 4 -1e-3\n2 2 7\n")
    nr, nc, i, j, v = mm.read_coo(p)
    assert (nr, nc) == (3, 4)
    assert i.tolist() == [0, 2, 1] and j.tolist() == [0, 3, 1]
    assert v.tolist() == [2.5, -1e-3, 7.0]


def test_symmetric_mirrors_off_diagonal_right_after(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 1\n3 1 2\n3 2 3\n")
    nr, nc, i, j, v = mm.read_coo(p)
    assert i.tolist() == [0, 2, 0, 2, 1] and j.tolist() == [0, 0, 2, 1, 2]
    assert v.tolist() == [1.0, 2.0, 2.0, 3.0, 3.0]


def test_skew_symmetric_negates_mirror(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 4\n")
    _, _, i, j, v = mm.read_coo(p)
    assert list(zip(i.tolist(), j.tolist(), v.tolist())) == [(1, 0, 4.0), (0, 1, -4.0)]


def test_pattern_and_integer(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 2\n2 1\n")
    assert mm.read_coo(p)[4].tolist() == [1.0, 1.0]
    p = write(tmp_path, "%%MatrixMarket matrix coordinate integer symmetric\n2 2 2\n1 1 -3\n2 1 5\n", "b.mtx")
    _, _, i, j, v = mm.read_coo(p)
    assert v.tolist() == [-3.0, 5.0, 5.0] and i.tolist() == [0, 1, 0]


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",          # not coordinate
    "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",    # complex
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",         # row out of range
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n",         # truncated
])
def test_rejects(tmp_path, text):
    with pytest.raises(Spc5Error):
        mm.read_coo(write(tmp_path, text))


def test_missing_file(tmp_path):
    with pytest.raises(Spc5Error):
        mm.read_coo(str(tmp_path / "none.mtx"))
