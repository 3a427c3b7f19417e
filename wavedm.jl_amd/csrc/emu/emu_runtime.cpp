// Fiber scheduler for the CPU emulation build (debug/test only; see emu_runtime.hpp).
#include "emu_runtime.hpp"

#include <cstdio>
#include <ucontext.h>
#include <vector>


namespace emu {

Ctx g_ctx;

namespace {
struct Fiber {
    ucontext_t uc;
    bool done = false;
    bool at_barrier = false;
    emu_dim3 tidx;
};

constexpr size_t kStack = 512 * 1024;
std::vector<Fiber> fibers;
std::vector<char*> stacks;
ucontext_t sched_uc;
int cur = -1;
const std::function<void()>* g_body = nullptr;

void switch_to_sched() {
    swapcontext(&fibers[cur].uc, &sched_uc);
}

void fiber_entry() {
    (*g_body)();
    fibers[cur].done = true;
    swapcontext(&fibers[cur].uc, &sched_uc);
}
}  // namespace

[[noreturn]] void fail(const char* msg) {
    std::fprintf(stderr, "wavedm-emu: %s\n", msg);
    std::abort();
}

void barrier() {
    if (cur < 0) fail("barrier outside a kernel");
    fibers[cur].at_barrier = true;
    switch_to_sched();
}

void launch(dim3 grid, dim3 block, const std::function<void()>& body) {
    const int nt = int(block.x * block.y * block.z);
    if (nt <= 0 || nt > 1024) fail("bad block size");
    if ((int)stacks.size() < nt) {
        for (int i = (int)stacks.size(); i < nt; ++i) stacks.push_back(new char[kStack]);
    }
    fibers.assign(nt, Fiber());
    g_body = &body;
    g_ctx.bdim = block;
    g_ctx.gdim = grid;
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                g_ctx.bid = emu_dim3(bx, by, bz);
                for (int t = 0; t < nt; ++t) {
                    Fiber& f = fibers[t];
                    f.done = false;
                    f.at_barrier = false;
                    f.tidx = emu_dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
                    getcontext(&f.uc);
                    f.uc.uc_stack.ss_sp = stacks[t];
                    f.uc.uc_stack.ss_size = kStack;
                    f.uc.uc_link = nullptr;
                    makecontext(&f.uc, fiber_entry, 0);
                }
                for (;;) {
                    for (int t = 0; t < nt; ++t) {
                        Fiber& f = fibers[t];
                        if (f.done) continue;
                        cur = t;
                        f.at_barrier = false;
                        g_ctx.tid = f.tidx;
                        swapcontext(&sched_uc, &f.uc);
                    }
                    cur = -1;
                    int ndone = 0, nbar = 0;
                    for (int t = 0; t < nt; ++t) {
                        ndone += fibers[t].done;
                        nbar += fibers[t].at_barrier;
                    }
                    if (ndone == nt) break;
                    if (ndone != 0) fail("barrier reached by only part of the block");
                    if (nbar != nt) fail("scheduler inconsistency");
                }
            }
    g_body = nullptr;
}

}  // namespace emu
