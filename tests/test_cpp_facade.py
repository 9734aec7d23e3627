"""The C++ facade include/hashdag_amd.hpp (the reference's include/hashdag names over the C-ABI).

CPU: the header compiles on its own with the reference's -fno-exceptions. GPU: tests/cpp/facade_check
(built by __graft_entry__.build()) drives Edit / ThreadedEdit / ThreadedGC / Traversal through the
facade and checks them word for word against direct C-ABI calls on a second pool."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_check")


def test_header_compiles_without_exceptions(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "hashdag_amd.hpp"\n'
                   'static_assert(hashdag_amd::DeviceEditor<hashdag_amd::AABBEditor>);\n'
                   'static_assert(!hashdag_amd::NodePointer::Null().HasValue());\n'
                   'int main() { return hashdag_amd::Config::Validate(hashdag_amd::DefaultConfig{}()) ? 0 : 1; }\n')
    r = subprocess.run(["g++", "-std=c++20", "-fno-exceptions", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_facade_matches_capi():
    assert os.path.exists(BIN), "tests/cpp/facade_check is missing: run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "facade ok" in r.stdout, r.stdout + r.stderr
