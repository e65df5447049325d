"""The command-line runner's argument handling (run_csr.cc:192-225) without a
GPU: compressed and directed inputs are refused as the reference refuses
them, and the flags parse to the reference's defaults."""
import pytest

from byo_amd import run as R


def test_defaults_follow_run_csr():
    a = R.parse(["-s", "g.adj"])
    assert (a.rounds, a.src, a.max_batch, a.i, a.b) == (3, 0, 1_000_000, False, False)
    a = R.parse(["-s", "-b", "-rounds", "5", "-src", "10", "-i", "-max_batch", "100", "--cpma", "g.adj"])
    assert (a.rounds, a.src, a.max_batch, a.i, a.b, a.cpma) == (5, 10, 100, True, True, True)


@pytest.mark.parametrize("argv,msg", [(["-s", "-c", "g.adj"], "reads in uncompressed files"),
                                      (["g.adj"], "does not support directed graphs yet"),
                                      (["-s"], "exactly one of")])
def test_refusals(argv, msg, capsys):
    assert R.main(argv) == -1
    assert msg in capsys.readouterr().err
