"""The host-side C++ sources (key parser, construct_Hal, dipoles, sample draws,
hierarchy and coupling tables, time grid, batch plan) and the C oracle, built with
the address and undefined-behaviour sanitizers (CPU only), run over the reference's
key files and over damaged copies of them: truncated at many points, and with each
number replaced by hostile values (negative, zero, huge, not a number).  Any sanitizer
finding aborts the driver, which fails the test; parse errors are expected and fine.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "eto-heom_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _build(cwd, target):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host compiler")
    r = subprocess.run(["make", "-s", target], cwd=cwd, capture_output=True, text=True)
    if r.returncode:
        if "asan" in r.stderr.lower() and "cannot find" in r.stderr.lower():
            pytest.skip("sanitizer runtime not installed")
        raise AssertionError(r.stderr)


def _template():
    txt = open(os.path.join(GOLDEN, "cpu_dimer_ohmic.key-tmpl")).read()
    for name, v in zip(("TAU1", "TAU2", "TAU3", "T_END"), (40, 90, 190, 300)):
        txt = txt.replace(name, str(v))
    return txt


def _inputs(tmp):
    base = {"tmpl": _template(), "dyn": open(os.path.join(GOLDEN, "key_0_0.key")).read()}
    paths = []

    def put(name, text):
        p = os.path.join(tmp, name)
        with open(p, "w") as f:
            f.write(text)
        paths.append(p)

    for b, txt in base.items():
        put(f"{b}.key", txt)
        for cut in range(0, len(txt), 23):
            put(f"{b}_cut{cut}.key", txt[:cut])
        nums = [m.span() for m in re.finditer(r"-?\d+(\.\d+)?([eE][-+]?\d+)?", txt)]
        for i, (s, e) in enumerate(nums):
            for j, bad in enumerate(("-3", "0", "100000000", "1e308", "nan", "x")):
                put(f"{b}_n{i}_{j}.key", txt[:s] + bad + txt[e:])
    put("empty.key", "")
    put("blank_keywords.key", "SIZE\n\n\nHEOM\n\n")
    return paths


def test_host_sources_under_sanitizers(tmp_path):
    _build(CSRC, "sanitize")
    exe = os.path.join(CSRC, "_build", "host_check_asan")
    paths = _inputs(str(tmp_path))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:allocator_may_return_null=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, *paths], capture_output=True, text=True, env=env, timeout=900)
    if r.returncode:
        done = {ln.split()[-2] for ln in r.stdout.splitlines() if ln.startswith(("ok ", "err "))}
        first = next((p for p in paths if p not in done), None)
        bad = open(first).read()[:400] if first else ""
        raise AssertionError(f"sanitizer finding on {first}:\n{bad}\n---\n{r.stderr[-3000:]}")
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("done"), last
    ok = int(last.split()[1])
    assert ok >= 4   # the unmodified keys and many damaged ones still parse and build


def test_oracle_under_sanitizers():
    _build(ORACLE, "sanitize")
    exe = os.path.join(ORACLE, "_build", "oracle_check_asan")
    env = dict(os.environ, UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "oracle_check ok" in r.stdout
