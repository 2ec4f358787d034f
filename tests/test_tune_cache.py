"""CPU: the process-wide matcher tuning cache of the batching layer (host
logic only, no GPU call): a file written by yume_pdq_amd_tune_save reads back
as the same entries, entries merge bucket by bucket, and a malformed file
loads nothing."""
import yume_pdq

FILE = """# comment lines and blank lines are skipped

gfx950:sramecc+:xnack- 256 31 2 1000000 5000000 : 8=valu 16=mfma:2 1024=mfma:8
gfx950:sramecc+:xnack- 256 31 1 10000000 : 8=valu 128=mfma:16
"""


def entries(text):
    return sorted(line for line in text.splitlines() if line and not line.startswith("#"))


def test_save_load_round_trip(tmp_path):
    yume_pdq.tune_clear()
    src = tmp_path / "tune.txt"
    src.write_text(FILE)
    assert yume_pdq.tune_load(str(src)) == 2
    out = tmp_path / "saved.txt"
    assert yume_pdq.tune_save(str(out)) == 2
    assert entries(out.read_text()) == entries(FILE)
    # loading the saved file again changes nothing
    assert yume_pdq.tune_load(str(out)) == 2
    again = tmp_path / "again.txt"
    yume_pdq.tune_save(str(again))
    assert again.read_text() == out.read_text()
    assert yume_pdq.tune_clear() == 2
    assert yume_pdq.tune_clear() == 0


def test_entries_merge_by_bucket(tmp_path):
    yume_pdq.tune_clear()
    a = tmp_path / "a.txt"
    a.write_text("gfx950 256 31 1 100 : 8=valu 64=mfma:4\n")
    b = tmp_path / "b.txt"
    b.write_text("gfx950 256 31 1 100 : 64=mfma:8 512=mfma:16\ngfx950 256 40 1 100 : 8=valu\n")
    assert yume_pdq.tune_load(str(a)) == 1
    assert yume_pdq.tune_load(str(b)) == 2
    out = tmp_path / "out.txt"
    assert yume_pdq.tune_save(str(out)) == 2
    assert entries(out.read_text()) == ["gfx950 256 31 1 100 : 8=valu 64=mfma:8 512=mfma:16",
                                        "gfx950 256 40 1 100 : 8=valu"]
    yume_pdq.tune_clear()


def test_malformed_files_load_nothing(tmp_path):
    yume_pdq.tune_clear()
    bad = ["gfx950 256 31 2 100 : 8=valu\n",                  # fewer row counts than shards
           "gfx950 256 31 1 100 8=valu\n",                    # no colon
           "gfx950 256 31 1 100 : 8=mfma:3\n",                # query tiles not 2/4/8/16
           "gfx950 256 31 1 100 : 0=valu\n",                  # bucket 0
           "gfx950 256 31 1 100 :\n",                         # no choices
           "gfx950 256 31 1 100 : 8=valu\ngfx950 x 31\n"]     # a good line, then a bad one
    for i, text in enumerate(bad):
        p = tmp_path / f"bad{i}.txt"
        p.write_text(text)
        try:
            yume_pdq.tune_load(str(p))
        except yume_pdq.YumePdqError:
            pass
        else:
            raise AssertionError(f"loaded {text!r}")
        assert yume_pdq.tune_clear() == 0, text
    try:
        yume_pdq.tune_load(str(tmp_path / "missing.txt"))
    except yume_pdq.YumePdqError:
        pass
    else:
        raise AssertionError("loaded a missing file")
