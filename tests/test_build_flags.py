"""The library's build lines and HIP sources, checked on the CPU: every unit
is compiled for gfx950 with -O3 and with packed FP32 instructions switched
off, without FMA contraction except the three units built for it (the mixed
path, the contracted FP64 path and the discrete adjoint); and no source asks
for a non-default stream priority, writes packed FP32 or any other machine
instruction by hand, registers existing host memory or allocates uncached
device memory."""
import os
import re

import pytest

from qasmhd_amd import build

CONTRACTED = {"qas_mixed.hip": "fast", "qas_fma64.hip": "fast", "qas_adjoint2d.hip": "on"}


@pytest.fixture
def lines(monkeypatch):
    seen = []

    class Done:
        returncode = 0
        stderr = ""

    def fake_run(cmd, **kw):
        seen.append(list(cmd))
        return Done()

    monkeypatch.delenv("QAS_OFFLOAD_ARCH", raising=False)
    monkeypatch.setattr(build.subprocess, "run", fake_run)
    for src in build.sources():
        build._compile(src)
    return dict(zip(build.sources(), seen))


def test_every_unit_is_gfx950_o3_without_packed_fp32(lines):
    assert len(lines) >= 10
    for src, cmd in lines.items():
        assert "--offload-arch=gfx950" in cmd, src
        assert "-O3" in cmd, src
        s = " ".join(cmd)
        assert "-Xclang -target-feature -Xclang -packed-fp32-ops" in s, src
        assert not any(f.startswith("-fsanitize") for f in cmd), src


def test_contraction_only_where_intended(lines):
    for src, cmd in lines.items():
        fc = [f for f in cmd if f.startswith("-ffp-contract=")]
        assert fc == ["-ffp-contract=" + CONTRACTED.get(src, "off")], (src, fc)
        fast_div = "-fno-hip-fp32-correctly-rounded-divide-sqrt" in cmd
        assert fast_div == (src == "qas_mixed.hip"), src
        assert ("-freciprocal-math" in cmd) == (src == "qas_adjoint2d.hip"), src


def test_sources_avoid_forbidden_constructs():
    bad = {
        "stream priority": re.compile(r"StreamCreateWithPriority|priority\s*="),
        "hand-written packed FP32": re.compile(r"v_pk_(add|mul|fma)_f32"),
        "registering existing host memory": re.compile(r"hipHostRegister"),
        "uncached device memory": re.compile(r"hipExtMallocWithFlags|hipDeviceMallocUncached"),
    }
    asm = re.compile(r"\basm\b[^;]*?\(\s*\"([^\"]*)\"")
    for src in os.listdir(build.CSRC):
        text = open(os.path.join(build.CSRC, src)).read()
        for what, rx in bad.items():
            assert not rx.search(text), (src, what)
        # the only inline asm allowed is an empty compiler barrier
        for body in asm.findall(text):
            assert body.strip() == "", (src, body)
