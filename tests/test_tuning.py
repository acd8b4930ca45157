"""The mixed mode's launch-shape table (am-cfd_amd/csrc/tuning.cpp): parsed
and looked up through the C ABI, without a GPU."""
import os

import pytest

from conftest import ROOT, amcfd_module


def write(tmp_path, text):
    p = tmp_path / "tuning.txt"
    p.write_text(text)
    return str(p)


def test_first_row_that_holds_the_grid_applies(tmp_path):
    amcfd = amcfd_module()
    path = write(tmp_path, "# max_cells graph_block box_blocks [box_fused [line_batch [overlap]]]\n"
                           "1000000 16 128 0 4 1   # small grids\n"
                           "\n"
                           "6000000 8 512\n")
    assert amcfd.tuning_choice(1, path) == (16, 128, False, 4, True)
    assert amcfd.tuning_choice(1000000, path) == (16, 128, False, 4, True)
    # box_fused and line_batch default to 1, the overlap to off
    assert amcfd.tuning_choice(1000001, path) == (8, 512, True, 1, False)
    assert amcfd.tuning_choice(6000000, path) == (8, 512, True, 1, False)
    assert amcfd.tuning_choice(6000001, path) == (8, 256, True, 1, False)   # past the last row: the built-in shapes


def test_empty_table_gives_the_built_in_shapes(tmp_path):
    amcfd = amcfd_module()
    empty = write(tmp_path, "# nothing tuned\n")
    for cells in (1, 300000, 5120000, 10 ** 9):
        assert amcfd.tuning_choice(cells, empty) == (8, 256, True, 1, False)


@pytest.mark.parametrize("bad", ["1000 8\n", "1000 0 256\n", "1000 8 5000\n", "0 8 256\n", "1000 8 256 7\n",
                                 "1000 8 256 1 0\n", "1000 8 256 1 65\n", "1000 8 256 1 2 3\n",
                                 "1000 8 256 1 2 1 4\n", "x 8 256\n"])
def test_malformed_rows_are_refused(tmp_path, bad):
    amcfd = amcfd_module()
    with pytest.raises(amcfd.AmcfdError) as e:
        amcfd.tuning_choice(10, write(tmp_path, "100 8 256\n" + bad))
    assert e.value.code == -3 and ":2:" in str(e.value)


def test_missing_table_file_is_an_io_error(tmp_path):
    amcfd = amcfd_module()
    with pytest.raises(amcfd.AmcfdError) as e:
        amcfd.tuning_choice(10, str(tmp_path / "absent.txt"))
    assert e.value.code == -2


def test_library_table_is_the_committed_one():
    """With no path the library reads tuning.txt beside libamcfd.so; the
    committed table's rows are what it returns."""
    amcfd = amcfd_module()
    table = os.path.join(ROOT, "am-cfd_amd", "tuning.txt")
    if os.environ.get("AMCFD_TUNING") or not os.path.exists(table):
        pytest.skip("no committed table, or a table chosen by the environment")
    for cells in (1, 300000, 1000000, 3000000, 5120000, 9700000, 10 ** 9):
        assert amcfd.tuning_choice(cells) == amcfd.tuning_choice(cells, table), cells
