"""The pipelined form's ticket layout (pipeline_task in
tpr_amd/csrc/launch_split.hpp, the function the kernel decodes its tickets
with), compiled for the host and enumerated over many round shapes and
lags: every rows, reduced and stage-3 task appears exactly once, and every
dependency -- all rows tasks of a system before its reduced task, the
reduced task before all of the system's stage-3 tasks -- has a lower ticket,
which is what makes the kernel's waits deadlock-free whatever the residency."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include <cstdio>
#include <vector>
#include "launch_split.hpp"
using namespace tprh;
int main() {
    long long rounds = 0;
    for (int lg = 0; lg <= 7; lg++) {
        const int bps = 1 << lg;
        for (int nsys : {1, 2, 3, 5, 8, 17, 70, 256}) {
            const int na = nsys * bps;
            for (int lag : {0, 1, bps, 2 * bps, 2 * bps + 1, 3 * bps + 1, 5 * bps, na - 1, na, na + 5, 768, 2048,
                            8192}) {
                const int D = pipeline_d(lag, na, bps);
                const int g = pipeline_gap(D, bps, nsys);
                const int ntk = pipeline_tickets(nsys, lg);
                std::vector<int> rows(na, -1), st3(na, -1), red(nsys, -1);
                for (int tk = 0; tk < ntk; tk++) {
                    const PipeTask t = pipeline_task(tk, nsys, lg, D, g);
                    std::vector<int>* v = t.kind == kPipeRows ? &rows : t.kind == kPipeStage3 ? &st3 : &red;
                    const int idx = t.kind == kPipeReduced ? t.i / bps : t.i;
                    if (t.kind == kPipeReduced && t.i % bps) { printf("reduced index %d\n", t.i); return 1; }
                    if (idx < 0 || idx >= (int)v->size() || (*v)[idx] >= 0) {
                        printf("bad or repeated task: kind %d i %d (nsys %d bps %d lag %d)\n", t.kind, t.i, nsys, bps,
                               lag);
                        return 1;
                    }
                    (*v)[idx] = tk;
                }
                for (int q = 0; q < nsys; q++) {
                    int last_rows = -1, first_st3 = ntk;
                    for (int b = 0; b < bps; b++) {
                        const int r = rows[q * bps + b], s = st3[q * bps + b];
                        if (r < 0 || s < 0) { printf("missing task\n"); return 1; }
                        last_rows = r > last_rows ? r : last_rows;
                        first_st3 = s < first_st3 ? s : first_st3;
                    }
                    if (!(last_rows < red[q] && red[q] < first_st3)) {
                        printf("order: nsys %d bps %d lag %d q %d: %d %d %d\n", nsys, bps, lag, q, last_rows, red[q],
                               first_st3);
                        return 1;
                    }
                }
                rounds++;
            }
        }
    }
    printf("%lld\n", rounds);
    return 0;
}
'''


def test_pipeline_tickets_cover_every_task_in_dependency_order(tmp_path):
    src = tmp_path / "tickets.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "tickets"
    p = subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I" + os.path.join(ROOT, "tpr_amd", "csrc"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert int(r.stdout.strip()) == 8 * 8 * 13
