"""The rocprofv3 CSV summariser on a small synthetic trace."""
import csv

from pardiso2_amd import rocprof_summary as R


def _trace(tmp_path, rows):
    with open(tmp_path / "x_kernel_trace.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Dispatch_Id"])
        w.writeheader()
        for i, (name, t0, t1) in enumerate(rows):
            w.writerow({"Kernel_Name": name, "Start_Timestamp": t0, "End_Timestamp": t1, "Dispatch_Id": i})


def test_window_from_the_kth_last_marker(tmp_path, capsys):
    _trace(tmp_path, [("setup", 0, 10), ("mark", 100, 110), ("a", 110, 150), ("mark", 1000, 1010),
                      ("a", 1010, 1050), ("b", 1100, 1200)])
    R.window(str(tmp_path), "mark", 1, 10)
    out = capsys.readouterr().out
    assert "window: 3 dispatches, span 0.000 ms" in out
    assert "(75.0 %)" in out  # (10 + 40 + 100) ns of a 200 ns span
    R.window(str(tmp_path), "mark", 2, 10)
    out = capsys.readouterr().out
    assert "window: 5 dispatches" in out
    # ended with the first later dispatch of "a": the last marker and a (50 ns of a 50 ns span)
    R.window(str(tmp_path), "mark", 1, 10, through="a")
    out = capsys.readouterr().out
    assert "window: 2 dispatches" in out
    assert "(100.0 %)" in out


def test_trace_last_dispatches(tmp_path, capsys):
    _trace(tmp_path, [("k1", 0, 1000), ("k1", 2000, 5000), ("k2", 0, 10)])
    R.trace(str(tmp_path), "k1", 1)
    out = capsys.readouterr().out
    assert "dispatches 1  avg us 3.000" in out


def test_window_by_launch_shape(tmp_path, capsys):
    cols = ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Dispatch_Id", "Grid_Size_X", "Workgroup_Size_X"]
    with open(tmp_path / "x_kernel_trace.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for i, (name, t0, t1, g) in enumerate([("mark", 0, 10, 64), ("sweep", 10, 30, 1024),
                                                ("sweep", 30, 40, 256), ("sweep", 40, 60, 1024)]):
            w.writerow(dict(zip(cols, [name, t0, t1, i, g, 64])))
    R.window(str(tmp_path), "mark", 1, 10, by_grid=True)
    out = capsys.readouterr().out
    assert "       2      0.000       0.02" in out and "[grid 1024/64] sweep" in out  # two launches, 20 ns each
    assert "[grid 256/64] sweep" in out
