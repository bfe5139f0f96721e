"""The request server's host bookkeeping (honei_amd/csrc/serve.hip) under
AddressSanitizer and UBSan on the CPU.

serve.hip is compiled host-only and linked against a simulated device
(tests/serve_host): streams are deferred queues that run when the host waits,
queries an event or at random moments; a pinned range is poisoned while a
queued copy owns it; slot memory is heap memory, so a tick that reaches a
freed or stale slot is a use-after-free or a contract failure; allocation and
synchronous copies during a capture abort.  tests/serve_host/driver.cc runs
seeded workloads through the C-ABI and checks every request's result, the
budget, the counters and per-level admission order.

test_harness_catches_planted_mistakes pins the harness itself: three copies of
serve.hip with one mistake each (writing the pinned table image while its
upload is queued, reading the reports before their event, not re-uploading
the table after retiring a slot) must each fail."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(ROOT, "tests", "serve_host")
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                                reason="needs hipcc and the ROCm clang with its sanitizer runtimes")


def _hip_host(src, obj):
    # host-only HIP compile: no device code is emitted, the sanitizers apply to the host pass
    subprocess.run([HIPCC, "--offload-host-only", "-x", "hip", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "honei_amd", "csrc"), "-I" + SIM,
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-c", src, "-o", obj],
                   check=True, capture_output=True)


def _cxx(src, obj):
    subprocess.run([CLANG, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-I" + SIM,
                    "-c", src, "-o", obj], check=True, capture_output=True)


def _link(objs, exe):
    subprocess.run([CLANG, "-fsanitize=address,undefined"] + objs + ["-o", exe], check=True, capture_output=True)


def _run(exe, seeds, timeout=240):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    return subprocess.run([exe] + [str(s) for s in seeds], capture_output=True, text=True, timeout=timeout, env=env)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("serve_host"))
    objs = {}
    for name, build in (("sim_mg", _hip_host), ("sim_hip", _cxx), ("driver", _cxx)):
        src = os.path.join(SIM, name + ".cc")
        objs[name] = os.path.join(d, name + ".o")
        build(src, objs[name])
    serve_o = os.path.join(d, "serve.o")
    _hip_host(os.path.join(ROOT, "honei_amd", "csrc", "serve.hip"), serve_o)
    exe = os.path.join(d, "driver")
    _link([serve_o, objs["sim_mg"], objs["sim_hip"], objs["driver"]], exe)
    return d, objs, exe


def test_serve_host_bookkeeping_under_asan(harness):
    _, _, exe = harness
    seeds = list(range(1, 601))
    r = _run(exe, seeds)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("ok ")]
    assert len(lines) == len(seeds)
    # every workload kind is reached
    for kind in ("trace", "destroyed in flight", "submit/step/poll", "graphs 1", "graphs 0", "budget mode 0"):
        assert any(kind in ln for ln in lines), kind


PLANTED = [
    ("table image written while its upload is queued",
     "        HIP_CHECK(hipEventRecord(b.ev, b.st));\n        b.inflight = true;",
     "        HIP_CHECK(hipEventRecord(b.ev, b.st));\n        b.table_host[0].partials = nullptr;\n"
     "        b.inflight = true;",
     "use-after-poison"),
    ("reports read before their event",
     "            if (e == hipErrorNotReady)\n                continue;",
     "            if (e == hipErrorNotReady && b.rep_host[0].phase == 7)\n                continue;\n"
     "            if (e == hipErrorNotReady)\n                continue;",
     "use-after-poison"),
    ("table not re-uploaded after a slot is retired",
     "            std::memset(&b.table_host[i], 0, sizeof(mg::Slot));\n            b.dirty = true;",
     "            std::memset(&b.table_host[i], 0, sizeof(mg::Slot));",
     "SIM CONTRACT"),
]


@pytest.mark.parametrize("what,old,new,sign", PLANTED, ids=[p[0] for p in PLANTED])
def test_harness_catches_planted_mistakes(harness, what, old, new, sign):
    d, objs, _ = harness
    with open(os.path.join(ROOT, "honei_amd", "csrc", "serve.hip")) as f:
        text = f.read()
    assert old in text, "the planted mistake no longer applies to serve.hip: " + what
    tag = str([x[0] for x in PLANTED].index(what))
    src = os.path.join(d, "planted_%s.hip" % tag)
    with open(src, "w") as f:
        f.write(text.replace(old, new, 1))
    obj = os.path.join(d, "planted_%s.o" % tag)
    _hip_host(src, obj)
    exe = os.path.join(d, "planted_%s" % tag)
    _link([obj, objs["sim_mg"], objs["sim_hip"], objs["driver"]], exe)
    r = _run(exe, range(1, 201))
    assert r.returncode != 0, what
    assert sign in r.stderr, (what, r.stderr[-3000:])
