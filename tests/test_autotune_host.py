"""Host logic of the autotuner: the tuning file's text format (written by
`write`, parsed by `read` and by the library at context creation) and the
selection rule for the comb width."""
from pathlib import Path

from ultrafastsecp256k1_amd import autotune as at

TUNING = Path(at.__file__).resolve().parent / "ufsecp_tuning.txt"


def test_shipped_tuning_file_round_trips(tmp_path):
    t = at.read(TUNING)
    assert {"gcomb_w", "verify_blocks_per_cu", "server_groups", "msm_window"} <= set(t)
    assert sorted(t["msm_window"]) == list(range(32))
    assert all(2 <= c <= 20 for c in t["msm_window"].values())
    out = tmp_path / "t.txt"
    at.write(t, out, TUNING.read_text().splitlines()[1].split(": ", 1)[1])
    assert out.read_text() == TUNING.read_text()


def test_read_keeps_values_and_record(tmp_path):
    t = {"gcomb_w": 14, "server_groups": 3, "msm_window": {0: 5, 20: 16},
         "_measured": {"server_groups": {"1": 10, "3": 12}}}
    out = tmp_path / "t.txt"
    at.write(t, out, "dev")
    assert at.read(out) == t


def test_smallest_within():
    assert at._smallest_within({12: 100.0, 14: 99.5, 16: 101.0}, 0.01) == 12
    assert at._smallest_within({12: 90.0, 14: 100.5, 16: 101.0}, 0.01) == 14
    assert at._smallest_within({12: 90.0, 14: 95.0, 16: 101.0}, 0.01) == 16
