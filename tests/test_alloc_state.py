"""All-or-nothing allocation of the fast-form operator buffers
(transportse_amd/csrc/tse_alloc.h, used by ensure_sys in tse_fast.hip): a
failure of the second allocation releases the first and leaves both
pointers null, so that a retry allocates again instead of finding a
half-built system.  Built with g++ against a stand-in allocator."""
import os
import subprocess

from helpers import ROOT

SRC = r"""
#include <cstdio>
#include <cstdlib>
#include "tse_alloc.h"
static int calls = 0, fail_at = -1, live = 0;
static bool fake_alloc(void **p, size_t n) {
    if (calls++ == fail_at) { *p = nullptr; return false; }
    *p = std::malloc(n); ++live; return true;
}
static void fake_free(void *p) { std::free(p); --live; }
int main() {
    for (int f = -1; f < 2; ++f) {
        calls = 0; fail_at = f;
        void *a = (void *)1, *b = (void *)1;
        bool ok = tse::alloc_both(&a, 64, &b, 32, fake_alloc, fake_free);
        if (f < 0) {
            if (!ok || !a || !b || live != 2) return 10;
            fake_free(a); fake_free(b);
        } else if (ok || a || b || live != 0) {
            return 20 + f;
        }
    }
    std::puts("ok");
    return 0;
}
"""


def test_failed_second_allocation_leaves_nothing(tmp_path):
    src = tmp_path / "alloc.cpp"
    src.write_text(SRC)
    exe = tmp_path / "alloc"
    inc = os.path.join(ROOT, "transportse_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", inc,
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.returncode
