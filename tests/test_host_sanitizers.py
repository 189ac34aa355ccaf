"""The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

The CPU suites of the ABI (argument validation before any HIP call), the Matrix Market
reader and the tuning tables run again in a child Python, against the host-sanitized
build of the library (`python -m segscan_amd.build --sanitize`, which __graft_entry__.build()
also makes), with the ASan runtime preloaded. Any sanitizer report fails the test
(UBSan is built with -fno-sanitize-recover, ASan halts on the first error).
"""
import subprocess
import sys
from pathlib import Path

import pytest

from segscan_amd import build

ROOT = Path(__file__).resolve().parent.parent
SUITES = ["tests/test_abi_cpu.py", "tests/test_mtx.py", "tests/test_tune_table.py"]


def test_host_code_clean_under_asan_and_ubsan():
    if not build.SAN_LIB.exists():
        pytest.skip(f"{build.SAN_LIB} not built (python -m segscan_amd.build --sanitize)")
    if build.LIB.exists() and build.SAN_LIB.stat().st_mtime < build.LIB.stat().st_mtime:
        pytest.skip(f"{build.SAN_LIB} is older than {build.LIB.name} (python -m segscan_amd.build --sanitize)")
    env = build.sanitized_env()
    env.pop("PYTEST_XDIST_WORKER", None)
    r = subprocess.run([sys.executable, "-m", "pytest", *SUITES, "-m", "not gpu", "-q", "-s", "-x",
                        "-p", "no:cacheprovider"], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    reports = [ln for ln in r.stderr.splitlines() if "runtime error" in ln or "Sanitizer" in ln]
    assert r.returncode == 0 and not reports, (r.stdout[-2000:] + "\n" + "\n".join(reports[:20]) + "\n" +
                                               r.stderr[-3000:])
    assert " passed" in r.stdout
