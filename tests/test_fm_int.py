"""Frame-integration tables (mc_fm_int_table) against hand-derived cases of
CReadFmIntFile / CFmIntParam (DataUtil/CReadFmIntFile.cpp:110-173,
DataUtil/CFmIntParam.cpp:118-262)."""
import numpy as np

from motioncor3_amd import lib as L


def _file(tmp_path, text):
    p = tmp_path / "fmint.txt"
    p.write_text(text)
    return str(p)


def test_no_file_is_one_raw_frame_each():
    t = L.fm_int_table(None, 10, 0.8, throw0=1, throw1=1)
    assert not t["integrated"]
    assert t["start"].tolist() == list(range(1, 9))
    assert t["size"].tolist() == [1] * 8
    # mSetup: FmDose x (i + 1), thrown frames not counted
    assert np.allclose(t["acc_dose"], 0.8 * np.arange(1, 9, dtype=np.float32))


def test_two_entries_exact(tmp_path):
    t = L.fm_int_table(_file(tmp_path, "20 2 0.5\n30 3 0.5\n"), 50, 9.0)
    assert t["integrated"] and t["raw_dose"] == 0.5
    assert t["size"].tolist() == [2] * 10 + [3] * 10
    assert t["start"].tolist() == list(range(0, 20, 2)) + list(range(20, 50, 3))
    # accumulated in float, frame by frame
    acc, a = [], np.float32(0)
    for s in t["size"]:
        a = np.float32(a + np.float32(s) * np.float32(0.5))
        acc.append(a)
    assert np.array_equal(t["acc_dose"], np.array(acc, np.float32))


def test_fewer_raw_frames_than_the_file(tmp_path):
    # 48 raw frames: the last entry loses 2 (28 = 9 x 3 + 1); the extra raw
    # frame goes to the last integrated frame
    t = L.fm_int_table(_file(tmp_path, "20 2 0.5\n30 3 0.5\n"), 48, 0.0)
    assert t["size"].tolist() == [2] * 10 + [3] * 8 + [4]
    assert int(t["size"].sum()) == 48


def test_more_raw_frames_than_the_file(tmp_path):
    t = L.fm_int_table(_file(tmp_path, "20 2 0.5\n30 3 0.5\n"), 53, 0.0)
    assert t["size"].tolist() == [2] * 10 + [3] * 11


def test_remainder_moves_to_the_next_entry(tmp_path):
    # 21 % 2 = 1 raw frame moves into the second entry: 31 = 10 x 3 + 1
    t = L.fm_int_table(_file(tmp_path, "21 2 1.0\n30 3 1.0\n"), 51, 0.0)
    assert t["size"].tolist() == [2] * 10 + [3] * 9 + [4]


def test_two_column_file_takes_the_flag_dose(tmp_path):
    t = L.fm_int_table(_file(tmp_path, "12 4\n"), 12, 1.25)
    assert t["integrated"] and t["raw_dose"] == 1.25 and t["size"].tolist() == [4, 4, 4]
    assert np.allclose(t["acc_dose"], [5.0, 10.0, 15.0])


def test_variable_dose_disables_integration(tmp_path):
    t = L.fm_int_table(_file(tmp_path, "20 2 0.5\n30 3 0.7\n"), 50, 0.6)
    assert not t["integrated"] and t["size"].tolist() == [1] * 50


def test_throws_keep_the_accumulated_dose(tmp_path):
    t = L.fm_int_table(_file(tmp_path, "40 4 0.25\n"), 40, 0.0, throw0=1, throw1=2)
    assert t["size"].tolist() == [4] * 7
    assert t["start"].tolist() == list(range(4, 32, 4))
    assert np.allclose(t["acc_dose"], [2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0])
    # fewer than six frames would remain: the throws are ignored
    t = L.fm_int_table(_file(tmp_path, "24 4 0.25\n"), 24, 0.0, throw0=1, throw1=0)
    assert t["size"].tolist() == [4] * 6


def test_unparsed_line_repeats_the_previous_entry(tmp_path):
    # the reference compares instead of assigning the sscanf count for
    # 3-column files, so the blank line repeats "10 2 1.0": entries 10 + 10
    # + 5 raw frames (without the repeat: 10 + 5, and the last entry would
    # absorb the 10 extra frames as [5] * 3)
    t = L.fm_int_table(_file(tmp_path, "10 2 1.0\n\n5 5 1.0\n"), 25, 0.0)
    assert t["size"].tolist() == [2] * 10 + [5]
