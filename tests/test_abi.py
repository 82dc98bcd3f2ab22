"""The C-ABI library loads and exports every function include/sherd_hip.h declares (CPU)."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sherd_hip.h")
LIB = os.path.join(ROOT, "sherd-refit_amd", "sherd_hip", "lib", "libsherd_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sherd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_executor_surface():
    names = declared()
    for must in ("sherd_open", "sherd_close", "sherd_bounded_distance", "sherd_inside", "sherd_first_hit",
                 "sherd_scene_build"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libsherd_hip.so not built (run sherd-refit_amd/build.py)")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sherd_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libsherd_hip.so not built")
def test_library_loads_without_a_device():
    import sherd_hip

    L = sherd_hip.lib()
    for n in declared():
        assert hasattr(L, n), n
    assert sherd_hip.HipExecutor.name() == "hip"
