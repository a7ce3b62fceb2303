"""The chunk and split-point arithmetic shared by the host plan and the
device plan (csrc/smoc_plan.h), compiled for the host with g++ and checked
over many track extents and chunk lengths: chunks tile each track in order,
none is empty unless the track is, a track of at most C segments stays
whole, chunks after a track's first are at most C long, and, for C a
multiple of the split length, every cut is a split point whose index maps
back to it (so the first run id of every chunk is a run-table entry)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "simplemoc-kernel_amd", "csrc")

HARNESS = r"""
#define __host__
#define __device__
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include "smoc_plan.h"

static int fail(const char *what, long long b, long long e, long long C, long long k)
{
    std::printf("FAIL %s b=%lld e=%lld C=%lld k=%lld\n", what, b, e, C, k);
    return 1;
}

int main()
{
    const int64_t L = SMOC_RUN_SPLIT;
    const int64_t Cs[] = {1, 7, 16, 100, 1000, 1023, 1024, 1025, 2048, 3072, 4096, 8192};
    uint64_t x = 12345;
    auto rnd = [&](int64_t n) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        return (int64_t)(x % (uint64_t)n);
    };
    long long checked = 0;
    for (int it = 0; it < 200000; ++it) {
        const int64_t b = rnd(60000);
        const int64_t lens[] = {0, 1, 2, L - 1, L, L + 1, rnd(40000)};
        const int64_t e = b + lens[rnd(7)];
        const int64_t C = Cs[rnd(12)];
        const int64_t nc = smoc_chunk_count(b, e, C);
        if (nc < 1) return fail("count", b, e, C, 0);
        if (e - b <= C && nc != 1) return fail("short track cut", b, e, C, 0);
        int64_t pos = b;
        for (int64_t k = 0; k < nc; ++k) {
            const int64_t cb = smoc_chunk_begin(b, k, C), ce = smoc_chunk_end(b, e, k, C);
            if (cb != pos) return fail("not contiguous", b, e, C, k);
            if (!(ce > cb || e == b)) return fail("empty chunk", b, e, C, k);
            if (k > 0 && ce - cb > C) return fail("chunk too long", b, e, C, k);
            if (k == 0 && ce - cb > C + (C % L == 0 ? L - 1 : 0))
                return fail("first chunk too long", b, e, C, k);
            if (k > 0 && C % L == 0) {
                if (cb % L != 0) return fail("cut off a split point", b, e, C, k);
                const int64_t j = smoc_split_index(b, cb);
                if (j < 1 || j > smoc_split_count(b, e)) return fail("split index", b, e, C, k);
                if ((b / L + j) * L != cb) return fail("split index maps back", b, e, C, k);
            }
            pos = ce;
        }
        if (pos != e) return fail("does not end at e", b, e, C, nc);
        int64_t ns = 0;
        for (int64_t p = (b / L + 1) * L; p < e; p += L)
            ++ns;
        if (smoc_split_count(b, e) != ns) return fail("split count", b, e, C, 0);
        ++checked;
    }
    std::printf("ok %lld\n", checked);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_chunk_and_split_arithmetic(tmp_path):
    src = tmp_path / "plan_harness.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "plan_harness"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, "-I", inc, str(src), "-o",
                    str(exe)], check=True, capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok 200000")
