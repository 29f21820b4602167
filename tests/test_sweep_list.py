"""Command lists in the reference's syntax (test/explist.bat): the runs, the
directories to create, defaults for omitted flags and relative paths."""
import os

from gahhg_amd import sweep


def test_parse_reference_syntax(tmp_path):
    lst = tmp_path / "explist.bat"
    lst.write_text(
        "md report\r\n"
        "md report\\GAHHG\\Mars\r\n"
        "\r\n"
        ".\\GPISP.exe /polyfile .\\polygon\\Mars\\001.txt /tpfile .\\query\\Mars_query.csv "
        "/reportfile .\\report\\GAHHG\\Mars\\report.txt /resultfile .\\report\\GAHHG\\Mars\\001_result.csv\r\n"
        ".\\GPISP.exe /polyfile .\\polygon\\various\\NA1.txt       /tpfile .\\query\\NA_query.csv   "
        "      /reportfile .\\report\\GAHHG\\various\\report.txt  \r\n"
        ".\\GPISP.exe /unknown x /polyfile /abs/a.txt\r\n"
        "rem not a run\r\n")
    jobs, dirs = sweep.parse_command_list(lst)
    t = str(tmp_path)
    assert dirs == [os.path.join(t, "report"), os.path.join(t, "report/GAHHG/Mars")]
    assert len(jobs) == 3
    assert jobs[0] == sweep.Job(os.path.join(t, "polygon/Mars/001.txt"), os.path.join(t, "query/Mars_query.csv"),
                                os.path.join(t, "report/GAHHG/Mars/report.txt"),
                                os.path.join(t, "report/GAHHG/Mars/001_result.csv"))
    # an omitted flag takes GPISPParam's default, relative to the list
    assert jobs[1].resultfile == os.path.join(t, "result.csv")
    assert jobs[1].tpfile == os.path.join(t, "query/NA_query.csv")
    # unknown flags are skipped; absolute paths stay as they are
    assert jobs[2].polyfile == "/abs/a.txt" and jobs[2].tpfile == os.path.join(t, "NA_query.csv")
    jobs, _ = sweep.parse_command_list(lst, base="/data")
    assert jobs[0].polyfile == "/data/polygon/Mars/001.txt"


def test_series_jobs():
    jobs = sweep.series_jobs(["a/001.txt", "a/002.txt"], "q.csv", "r.txt", "out")
    assert [j.resultfile for j in jobs] == ["out/001_result.csv", "out/002_result.csv"]
    assert all(j.tpfile == "q.csv" and j.reportfile == "r.txt" for j in jobs)
