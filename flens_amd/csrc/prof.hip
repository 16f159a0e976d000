// prof.hip -- launch profiler: HIP events around profiled launches, on the
// stream each launch uses, aggregated per kernel class with the algorithmic
// flops and bytes the caller attributes to the launch.
#include <vector>
#include "flx_kernels.h"

namespace flx {

struct ProfRec {
    hipEvent_t a, b;
    int cls;
    double flops, bytes;
};

struct Profiler {
    unsigned mask = ~0u;   // classes recorded (bit cls)
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
    size_t next = 0;
    hipEvent_t get()
    {
        if (next == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[next++];
    }
};

thread_local Profiler *tl_prof = nullptr;

Profiler *prof_new() { return new Profiler(); }

void prof_delete(Profiler *p)
{
    if (!p) return;
    for (hipEvent_t e : p->pool) (void)hipEventDestroy(e);
    delete p;
}

void prof_reset(Profiler *p)
{
    p->recs.clear();
    p->next = 0;
}

void prof_set_mask(Profiler *p, unsigned mask) { if (p) p->mask = mask; }

void *prof_begin(Profiler *p, hipStream_t s, int cls, double flops, double bytes)
{
    if (!p || !((p->mask >> cls) & 1u)) return nullptr;
    ProfRec r;
    r.a = p->get();
    r.b = p->get();
    if (!r.a || !r.b) return nullptr;
    r.cls = cls;
    r.flops = flops;
    r.bytes = bytes;
    (void)hipEventRecord(r.a, s);
    p->recs.push_back(r);
    return (void *)(uintptr_t)p->recs.size();   // 1-based index
}

void prof_end(Profiler *p, hipStream_t s, void *rec)
{
    size_t i = (size_t)(uintptr_t)rec - 1;
    (void)hipEventRecord(p->recs[i].b, s);
}

// out[cls*4 + {0,1,2,3}] = {launches, milliseconds, flops, bytes}
void prof_collect(Profiler *p, double *out)
{
    for (int i = 0; i < PROF_NCLS * 4; ++i) out[i] = 0;
    for (const ProfRec &r : p->recs) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        out[r.cls * 4 + 0] += 1;
        out[r.cls * 4 + 1] += ms;
        out[r.cls * 4 + 2] += r.flops;
        out[r.cls * 4 + 3] += r.bytes;
    }
}

} // namespace flx
