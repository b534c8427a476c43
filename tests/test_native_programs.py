"""Test programs built by __graft_entry__.build():

- tests/native/sancheck: the block codec's own source and the oracle built
  together under AddressSanitizer and UndefinedBehaviorSanitizer, agreeing bit
  for bit over seeded fields (specials included), every type, 1-D to 3-D,
  fixed and variable rate, the LDS phases run thread by thread, and random
  streams;
- tests/c/dropin: a C program written against include/zfp.h and linked to
  libzfp_hip.so, as a C caller of zfp-rs-ffi would be
  (zfp-rs-ffi/tests/c-dropin plays this part in the reference).  Without a GPU
  it runs the calls that need none; with one, compress and decompress too.
"""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
SANCHECK = ROOT / "tests" / "native" / "sancheck"
DROPIN = ROOT / "tests" / "c" / "dropin"


def run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout + p.stderr


def test_codec_and_oracle_under_sanitizers():
    if not SANCHECK.exists():
        pytest.fail(f"{SANCHECK} is not built: run __graft_entry__.build()")
    rc, out = run([str(SANCHECK)])
    assert rc == 0, out
    assert "cases agree" in out and "runtime error" not in out and "AddressSanitizer" not in out


def test_c_dropin_host_calls():
    if not DROPIN.exists():
        pytest.fail(f"{DROPIN} is not built: run __graft_entry__.build()")
    rc, out = run([str(DROPIN), "--host"])
    assert rc == 0, out
    assert "all checks passed (host)" in out


@pytest.mark.gpu
def test_c_dropin_on_the_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    rc, out = run([str(DROPIN)])
    assert rc == 0, out
    assert "all checks passed (host and GPU)" in out
