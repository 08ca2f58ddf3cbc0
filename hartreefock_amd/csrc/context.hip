// libhfgpu context: molecule upload, shell-pair data, Schwarz screening, task lists,
// and the ERI / J-K entry points of the C ABI (include/hfgpu.h).
#include <algorithm>
#include <limits>
#include <cmath>
#include <cstring>
#include <numeric>

#include "generated/dispatch.inc"
#include "hf_internal.h"

using namespace hfgpu;

namespace {

__global__ void fill_kernel(double* __restrict__ p, size_t n, double v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Deterministic mode: the fixed-point accumulator pairs fx[2 i], fx[2 i + 1] folded into out[i]
__global__ void fx_finalize_kernel(const long long* __restrict__ fx, double* __restrict__ out, size_t len) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x)
        out[i] = fx_value(fx[2 * i], fx[2 * i + 1]);
}

// out = in + in^T (per matrix), or out += in + in^T
__global__ void symmetrize_kernel(const double* __restrict__ in, double* __restrict__ out, int n, int count,
                                  int accumulate) {
    const size_t nn = (size_t)n * n;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn * count;
         idx += (size_t)gridDim.x * blockDim.x) {
        const size_t m = idx / nn, r = idx % nn;
        const size_t i = r / n, j = r % n;
        const double v = in[m * nn + i * n + j] + in[m * nn + j * n + i];
        out[idx] = accumulate ? out[idx] + v : v;
    }
}

int grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (int)std::min<size_t>(std::max<size_t>(b, 1), 8192);
}

}  // namespace

static int stage_reserve(hf_ctx* ctx, size_t need) {
    if (ctx->stage_used + need <= ctx->stage_cap) return HF_OK;
    // the queued copies still read the staged bytes: wait for them, then reuse (or grow) the buffer
    HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stage_used = 0;
    if (need > ctx->stage_cap) {
        const size_t cap = std::max<size_t>(need, std::max<size_t>(2 * ctx->stage_cap, (size_t)4 << 20));
        if (ctx->stage) (void)hipHostFree(ctx->stage);
        ctx->stage = nullptr;
        ctx->stage_cap = 0;
        HF_HIP_CHECK(ctx, hipHostMalloc((void**)&ctx->stage, cap, hipHostMallocDefault));
        ctx->stage_cap = cap;
    }
    return HF_OK;
}

int hf_upload(hf_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return HF_OK;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (int rc = stage_reserve(ctx, need)) return rc;
    char* p = ctx->stage + ctx->stage_used;
    std::memcpy(p, src, bytes);
    HF_HIP_CHECK(ctx, hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, ctx->stream));
    ctx->stage_used += need;
    return HF_OK;
}

void* hf_stage_host(hf_ctx* ctx, size_t bytes) {
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (stage_reserve(ctx, need)) return nullptr;
    void* p = ctx->stage + ctx->stage_used;
    ctx->stage_used += need;
    return p;
}

namespace {

int pair_class_index(int la, int lb) {
    for (int i = 0; i < kNumPairClasses; ++i)
        if (kPairClassL[i][0] == la && kPairClassL[i][1] == lb) return i;
    return -1;
}

int upload_pairs(hf_ctx* ctx, PairClassHost& pc) {
    const size_t np = pc.sa.size();
    HF_HIP_CHECK(ctx, pc.d_sa.alloc(np));
    HF_HIP_CHECK(ctx, pc.d_sb.alloc(np));
    HF_HIP_CHECK(ctx, pc.d_prim0.alloc(np));
    HF_HIP_CHECK(ctx, pc.d_nprim.alloc(np));
    HF_HIP_CHECK(ctx, pc.d_Q.alloc(np));
    if (!np) return HF_OK;
    if (int rc = hf_upload(ctx, pc.d_sa.p, pc.sa.data(), np * sizeof(int))) return rc;
    if (int rc = hf_upload(ctx, pc.d_sb.p, pc.sb.data(), np * sizeof(int))) return rc;
    if (int rc = hf_upload(ctx, pc.d_prim0.p, pc.prim0.data(), np * sizeof(int))) return rc;
    if (int rc = hf_upload(ctx, pc.d_nprim.p, pc.nprim.data(), np * sizeof(int))) return rc;
    if (int rc = hf_upload(ctx, pc.d_Q.p, pc.Q.data(), np * sizeof(double))) return rc;
    std::vector<PairRec> rec(np);
    for (size_t i = 0; i < np; ++i) {
        rec[i].sa = pc.sa[i];
        rec[i].sb = pc.sb[i];
        rec[i].prim0 = pc.prim0[i];
        rec[i].nprim = pc.nprim[i];
        rec[i].a0 = ctx->mol.shells[pc.sa[i]].bf;
        rec[i].b0 = ctx->mol.shells[pc.sb[i]].bf;
        rec[i].pad0 = rec[i].pad1 = 0;
    }
    HF_HIP_CHECK(ctx, pc.d_rec.alloc(np));
    if (int rc = hf_upload(ctx, pc.d_rec.p, rec.data(), np * sizeof(PairRec))) return rc;
    if (!pc.group.empty()) {
        HF_HIP_CHECK(ctx, pc.d_group.alloc(pc.group.size()));
        if (int rc = hf_upload(ctx, pc.d_group.p, pc.group.data(), pc.group.size() * sizeof(int))) return rc;
    }
    return HF_OK;
}

PairClassDev dev_view(const PairClassHost& pc) {
    PairClassDev d;
    d.rec = pc.d_rec.p;
    d.sa = pc.d_sa.p;
    d.sb = pc.d_sb.p;
    d.prim0 = pc.d_prim0.p;
    d.nprim = pc.d_nprim.p;
    d.Q = pc.d_Q.p;
    d.npairs = pc.npairs();
    return d;
}

// Statistics slot of a class launch: FP64 launches use [0, kNumClasses), FP32 ones the next kNumClasses.
int stat_slot(int mode, int cls) { return mode == MODE_JK32 ? kNumClasses + cls : cls; }

void add_stats(hf_ctx* ctx, int mode, const QuartetClassTasks& qc, const SinkArgs& sink) {
    const ClassInfo& ci = kClasses[qc.cls];
    const int pf = ctx->opt.exact_transfer && ci.exact_differs ? ci.prim_flops_exact : ci.prim_flops_ref;
    const int L = ci.la + ci.lb + ci.lc + ci.ld;
    // per primitive quartet: Boys (8-term Taylor 22, recursion 4/order, scaling 1/order)
    // and the quartet geometry/prefactor terms (36)
    const double boys_flops = 22 + 5.0 * L + 1 + 36;
    const int na = (ci.la + 1) * (ci.la + 2) / 2, nb = (ci.lb + 1) * (ci.lb + 2) / 2;
    const int nc = (ci.lc + 1) * (ci.lc + 2) / 2, nd = (ci.ld + 1) * (ci.ld + 2) / 2;
    const double nout = (double)na * nb * nc * nd;
    double sink_flops = 0;
    if (mode == MODE_JK || mode == MODE_JK32)
        sink_flops = nout * 2.0 * (2 + 4 * sink.ndens);   // 2 J + 4 K contractions per density
    KernelStat& ks = ctx->stats[stat_slot(mode, qc.cls)];
    ks.launches += 1;
    ks.quartets += qc.quartets;
    ks.prim_quartets += qc.prim_quartets;
    ks.eris += (double)qc.quartets * nout;
    ks.flops += (double)qc.prim_quartets * (pf + boys_flops) + (double)qc.quartets * (ci.hrr_flops + sink_flops);
}

TaskList task_list(hf_ctx* ctx, const TaskSet& ts, const QuartetClassTasks& qc) {
    TaskList tl{};
    tl.bra = dev_view(ctx->pcls[qc.bra]);
    tl.ket = dev_view(ctx->pcls[qc.ket]);
    tl.tasks = ts.d_tasks.p + qc.tasks_off;
    tl.prefix = ts.d_prefix.p;
    tl.koff = ts.d_koff.p;
    tl.kbeg = ts.d_kbeg.p;
    tl.kgroup = ctx->pcls[qc.ket].d_group.p;
    tl.seg = ts.d_seg.p + qc.seg_off;
    tl.nseg = qc.nseg;
    tl.ntasks = qc.ntasks;
    tl.nblocks = qc.nblocks;
    tl.same_class = qc.bra == qc.ket;
    tl.slice_prims = qc.slice_prims;
    return tl;
}

// Adds the HIP-event times of the last profiled build's launches to the statistics.  Called when
// the next build starts (by then the SCF step has long waited for that build) and when the
// statistics are read, so profiling adds no synchronisation to a J/K build.
int flush_kernel_times(hf_ctx* ctx) {
    for (int c : ctx->ev_pending) {
        HF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[2 * c + 1]));
        float ms = 0;
        HF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev[2 * c], ctx->ev[2 * c + 1]));
        ctx->stats[c].ms += ms;
    }
    ctx->ev_pending.clear();
    if (ctx->incore_pending >= 0) {
        HF_HIP_CHECK(ctx, hipEventSynchronize(ctx->incore_ev[1]));
        float ms = 0;
        HF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->incore_ev[0], ctx->incore_ev[1]));
        ctx->incore_stat[ctx->incore_pending].ms += ms;
        ctx->incore_pending = -1;
    }
    return HF_OK;
}

// Launch every quartet class of the given (mode, task set) jobs on the context's side streams
// between one fork from and one join into st; with record, add to the per-class statistics.
int run_classes(hf_ctx* ctx, const std::vector<std::pair<int, const TaskSet*>>& jobs, const SinkArgs& sink,
                hipStream_t st, bool record) {
    ShellsDev sh{ctx->d_shell_xyz.p, ctx->d_shell_bf.p};
    const int nslot = 2 * kNumClasses;
    if (ctx->stats.size() != (size_t)nslot) {
        ctx->stats.assign(nslot, KernelStat{});
        for (int c = 0; c < nslot; ++c) {
            ctx->stats[c].cls = c % kNumClasses;
            ctx->stats[c].fp32 = c >= kNumClasses;
        }
    }
    const bool prof = ctx->opt.profile != 0 && record;
    if (int rc = flush_kernel_times(ctx)) return rc;
    if (prof && ctx->ev.size() < 2 * (size_t)nslot) {
        for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
        ctx->ev.assign(2 * nslot, nullptr);
        for (auto& e : ctx->ev) HF_HIP_CHECK(ctx, hipEventCreate(&e));
    }
    // Only tiny builds (under 2^14 tasks) go in order on st: for them the fork/join across side
    // streams (~50 us) costs more than the overlap gains.  A small molecule's classes are each a
    // latency-bound handful of blocks, so running them side by side shortens its build (6-31G
    // waters and alkanes: SCF iterations -9 %); in a busy server the contexts' own overlap already
    // fills the chip, and its throughput is unchanged either way.
    int64_t total_tasks = 0;
    for (const auto& job : jobs)
        for (const QuartetClassTasks& qc : job.second->classes) total_tasks += qc.ntasks;
    constexpr int64_t kConcurrentMinTasks = 1 << 14;
    const bool fork = total_tasks >= kConcurrentMinTasks;
    // fork: every side stream waits for the work already queued on st
    constexpr int kSide = 4;
    if (fork && ctx->side.empty()) {
        ctx->side.assign(kSide, nullptr);
        for (auto& s2 : ctx->side) HF_HIP_CHECK(ctx, hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        ctx->side_ev.assign(kSide + 1, nullptr);
        for (auto& e : ctx->side_ev) HF_HIP_CHECK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if (fork) {
        HF_HIP_CHECK(ctx, hipEventRecord(ctx->side_ev[0], st));
        for (hipStream_t s2 : ctx->side) HF_HIP_CHECK(ctx, hipStreamWaitEvent(s2, ctx->side_ev[0], 0));
    }
    // Every class launch runs concurrently with the others (the high-L classes use the
    // one-wavefront-per-quartet LDS kernel, so no launch needs much scratch): largest estimated
    // work first, each to the least-loaded side stream.
    struct Launch {
        int mode;
        const TaskSet* ts;
        const QuartetClassTasks* qc;
        double cost;
    };
    std::vector<Launch> order;
    for (const auto& job : jobs)
        for (const QuartetClassTasks& qc : job.second->classes) {
            if (qc.ntasks <= 0) continue;
            const ClassInfo& ci = kClasses[qc.cls];
            // estimated duration x aggregate rate: the launch's flops plus the serial work of one quartet
            // times the ratio of the chip's aggregate to a single lane's dependent-FP64 rate (~2e4), so
            // that the latency-bound classes (few, long quartets) start first and overlap the others
            const double work = (double)qc.prim_quartets * (ci.prim_flops_ref + 60) + (double)qc.ntasks * ci.hrr_flops;
            const double chain = work / (double)qc.ntasks;
            // a class timed alone (profile_class) goes first, so its events bracket the kernel itself
            // rather than its wait for CUs held by the other streams' kernels
            const int key = ci.la * 1000 + ci.lb * 100 + ci.lc * 10 + ci.ld;
            const bool first = prof && ctx->opt.profile_class == key;
            order.push_back({job.first, job.second, &qc, first ? 1e300 : work + 2e4 * chain});
        }
    std::stable_sort(order.begin(), order.end(), [](const Launch& a, const Launch& b) { return a.cost > b.cost; });
    std::vector<double> load(kSide, 0.0);
    std::vector<int> launched;
    for (const Launch& L : order) {
        const QuartetClassTasks& qc = *L.qc;
        const ClassInfo& ci = kClasses[qc.cls];
        const int si = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[si] += L.cost;
        hipStream_t ks = fork ? ctx->side[si] : st;
        const TaskList tl = task_list(ctx, *L.ts, qc);
        const int slot = stat_slot(L.mode, qc.cls);
        const int key = ci.la * 1000 + ci.lb * 100 + ci.lc * 10 + ci.ld;
        const bool timed = prof && (ctx->opt.profile_class == -1 || ctx->opt.profile_class == key);
        if (timed) HF_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2 * slot], ks));
        hipError_t e = ci.fn(L.mode, ctx->opt.exact_transfer, tl, sh, ctx->d_pp.p, ctx->d_boys.p, sink, ks);
        if (e != hipSuccess) return ctx->fail(HF_E_HIP, std::string("ERI kernel launch: ") + hipGetErrorString(e));
        if (timed) {
            HF_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2 * slot + 1], ks));
            launched.push_back(slot);
        }
        if (record) add_stats(ctx, L.mode, qc, sink);
    }
    // join: st continues after every side stream's work
    for (int k = 0; fork && k < kSide; ++k) {
        HF_HIP_CHECK(ctx, hipEventRecord(ctx->side_ev[k + 1], ctx->side[k]));
        HF_HIP_CHECK(ctx, hipStreamWaitEvent(st, ctx->side_ev[k + 1], 0));
    }
    if (prof) ctx->ev_pending = launched;
    return HF_OK;
}

int run_classes(hf_ctx* ctx, int mode, const TaskSet& ts, const SinkArgs& sink, hipStream_t st, bool record) {
    return run_classes(ctx, {{mode, &ts}}, sink, st, record);
}

// Shell pairs, primitive pairs, Schwarz bounds, sorted pair lists.
int setup_molecule(hf_ctx* ctx) {
    hf_scf_session_free(ctx);
    const hfhost::Molecule& mol = ctx->mol;
    const int ns = (int)mol.shells.size();
    // the device buffers of the previous molecule are reused: its work must be finished
    HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->eri_stream) HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->eri_stream));
    for (hipStream_t s2 : ctx->side) HF_HIP_CHECK(ctx, hipStreamSynchronize(s2));
    ctx->stage_used = 0;                 // no staged upload is queued any more
    if (mol.maxL > kLmax) return ctx->fail(HF_E_UNSUPPORTED, "basis has L > 2 shells; this build supports s, p, d");
    std::vector<double> sxyz(3 * ns);
    std::vector<int> sbf(ns);
    for (int s = 0; s < ns; ++s) {
        for (int k = 0; k < 3; ++k) sxyz[3 * s + k] = mol.shells[s].xyz[k];
        sbf[s] = mol.shells[s].bf;
    }
    HF_HIP_CHECK(ctx, ctx->d_shell_xyz.alloc(sxyz.size()));
    HF_HIP_CHECK(ctx, ctx->d_shell_bf.alloc(sbf.size()));
    if (int rc = hf_upload(ctx, ctx->d_shell_xyz.p, sxyz.data(), sxyz.size() * sizeof(double))) return rc;
    if (int rc = hf_upload(ctx, ctx->d_shell_bf.p, sbf.data(), sbf.size() * sizeof(int))) return rc;
    if (!ctx->d_boys.p) {
        std::vector<double> tab;
        hfhost::boys_table(tab);
        HF_HIP_CHECK(ctx, ctx->d_boys.alloc(tab.size()));
        HF_HIP_CHECK(ctx, hipMemcpy(ctx->d_boys.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
        std::vector<float> tab32(tab.begin(), tab.end());
        HF_HIP_CHECK(ctx, ctx->d_boys32.alloc(tab32.size()));
        HF_HIP_CHECK(ctx, hipMemcpy(ctx->d_boys32.p, tab32.data(), tab32.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    // pairs (La >= Lb; for equal L, first shell index >= second)
    std::vector<PrimPair> pp;
    // pair classes keep their device buffers from the previous molecule (they only grow)
    ctx->pcls.resize(kNumPairClasses);
    for (PairClassHost& pc : ctx->pcls) {
        pc.sa.clear();
        pc.sb.clear();
        pc.prim0.clear();
        pc.nprim.clear();
        pc.Q.clear();
        pc.group.clear();
    }
    for (int sa = 0; sa < ns; ++sa)
        for (int sb = 0; sb < ns; ++sb) {
            const hfhost::Shell& A = mol.shells[sa];
            const hfhost::Shell& B = mol.shells[sb];
            if (A.L < B.L || (A.L == B.L && sb > sa)) continue;
            const int ci = pair_class_index(A.L, B.L);
            const double ABx = A.xyz[0] - B.xyz[0], ABy = A.xyz[1] - B.xyz[1], ABz = A.xyz[2] - B.xyz[2];
            const double R2 = ABx * ABx + ABy * ABy + ABz * ABz;
            const int start = (int)pp.size();
            for (size_t i = 0; i < A.alpha.size(); ++i)
                for (size_t j = 0; j < B.alpha.size(); ++j) {
                    const double a = A.alpha[i], b = B.alpha[j], z = a + b;
                    const double K = A.coef[i] * B.coef[j] * std::exp(-a * b / z * R2);
                    if (std::fabs(K) < ctx->opt.prim_threshold) continue;
                    PrimPair q;
                    q.Px = (a * A.xyz[0] + b * B.xyz[0]) / z;
                    q.Py = (a * A.xyz[1] + b * B.xyz[1]) / z;
                    q.Pz = (a * A.xyz[2] + b * B.xyz[2]) / z;
                    q.zeta = z;
                    q.Kz = K / z;
                    q.inv_zeta = 1.0 / z;
                    q.pad0 = q.pad1 = 0;
                    pp.push_back(q);
                }
            const int cnt = (int)pp.size() - start;
            if (cnt == 0) continue;
            PairClassHost& pc = ctx->pcls[ci];
            pc.sa.push_back(sa);
            pc.sb.push_back(sb);
            pc.prim0.push_back(start);
            pc.nprim.push_back(cnt);
            pc.Q.push_back(0.0);
        }
    HF_HIP_CHECK(ctx, ctx->d_pp.alloc(std::max<size_t>(pp.size(), 1)));
    if (!pp.empty())
        if (int rc = hf_upload(ctx, ctx->d_pp.p, pp.data(), pp.size() * sizeof(PrimPair))) return rc;
    for (auto& pc : ctx->pcls) {
        int rc = upload_pairs(ctx, pc);
        if (rc) return rc;
    }
    // Schwarz bounds from the diagonal quartets (ab|ab), always with exact integrals: every pair class's kernel,
    // then one read-back of all the bounds
    {
        size_t tot = 0;
        std::vector<size_t> off(ctx->pcls.size(), 0);
        for (size_t c = 0; c < ctx->pcls.size(); ++c) {
            off[c] = tot;
            tot += ctx->pcls[c].sa.size();
        }
        HF_HIP_CHECK(ctx, ctx->d_Qdiag.alloc(std::max<size_t>(tot, 1)));
        ShellsDev sh{ctx->d_shell_xyz.p, ctx->d_shell_bf.p};
        for (int c = 0; c < kNumClasses; ++c) {
            const ClassInfo& ci = kClasses[c];
            if (ci.la != ci.lc || ci.lb != ci.ld) continue;
            const int pc_i = pair_class_index(ci.la, ci.lb);
            PairClassHost& pc = ctx->pcls[pc_i];
            if (pc.sa.empty()) continue;
            TaskList tl{};
            tl.bra = tl.ket = dev_view(pc);
            tl.ntasks = pc.npairs();
            tl.same_class = 1;
            SinkArgs sink{};
            sink.Qout = ctx->d_Qdiag.p + off[pc_i];
            hipError_t e = ci.fn(MODE_DIAG, 1, tl, sh, ctx->d_pp.p, ctx->d_boys.p, sink, ctx->stream);
            if (e != hipSuccess) return ctx->fail(HF_E_HIP, std::string("Schwarz kernel: ") + hipGetErrorString(e));
        }
        if (tot) {
            double* hq = (double*)hf_stage_host(ctx, tot * sizeof(double));
            if (!hq) return ctx->fail(HF_E_OOM, "pinned staging buffer");
            HF_HIP_CHECK(ctx, hipMemcpyAsync(hq, ctx->d_Qdiag.p, tot * sizeof(double), hipMemcpyDeviceToHost,
                                             ctx->stream));
            HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
            for (size_t c = 0; c < ctx->pcls.size(); ++c)
                std::copy(hq + off[c], hq + off[c] + ctx->pcls[c].Q.size(), ctx->pcls[c].Q.begin());
        }
    }
    // sort each class into contraction groups (primitive-pair count, descending), each by Q, descending
    for (auto& pc : ctx->pcls) {
        const size_t np = pc.sa.size();
        std::vector<size_t> order(np);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
            if (pc.nprim[x] != pc.nprim[y]) return pc.nprim[x] > pc.nprim[y];
            return pc.Q[x] > pc.Q[y];
        });
        PairClassHost s;
        for (size_t i : order) {
            s.sa.push_back(pc.sa[i]);
            s.sb.push_back(pc.sb[i]);
            s.prim0.push_back(pc.prim0[i]);
            s.nprim.push_back(pc.nprim[i]);
            s.Q.push_back(pc.Q[i]);
        }
        pc.sa.swap(s.sa);
        pc.sb.swap(s.sb);
        pc.prim0.swap(s.prim0);
        pc.nprim.swap(s.nprim);
        pc.Q.swap(s.Q);
        pc.group.clear();
        for (size_t i = 0; i < np; ++i)
            if (i == 0 || pc.nprim[i] != pc.nprim[i - 1]) pc.group.push_back((int)i);
        pc.group.push_back((int)np);
        int rc = upload_pairs(ctx, pc);
        if (rc) return rc;
    }
    // per padded pair position, the Schwarz factor of its shell pair rounded up to FP32 (the in-core contraction
    // leaves out the loads of store elements the screening zeroed)
    {
        std::vector<double> qs((size_t)ns * ns, 0.0);
        for (const PairClassHost& pc : ctx->pcls)
            for (size_t q = 0; q < pc.sa.size(); ++q)
                qs[(size_t)pc.sa[q] * ns + pc.sb[q]] = qs[(size_t)pc.sb[q] * ns + pc.sa[q]] = pc.Q[q];
        const int n = mol.nbf;
        std::vector<int> shell_of(n, 0);
        for (int sh = 0; sh < ns; ++sh)
            for (int c = 0; c < hfhost::ncart(mol.shells[sh].L); ++c) shell_of[mol.shells[sh].bf + c] = sh;
        const int64_t np = pair_pad_base(n);
        std::vector<float> qf((size_t)np, 0.0f);
        double qmax = 0;
        for (int k = 0; k < n; ++k)
            for (int l = 0; l <= k; ++l) {
                const double q = qs[(size_t)shell_of[k] * ns + shell_of[l]];
                float f = (float)q;
                if ((double)f < q) f = std::nextafter(f, std::numeric_limits<float>::infinity());
                qf[(size_t)(pair_pad_base(k) + l)] = f;
                qmax = std::max(qmax, (double)f);
            }
        HF_HIP_CHECK(ctx, ctx->d_qf.alloc(std::max<size_t>(qf.size(), 1)));
        if (int rc = hf_upload(ctx, ctx->d_qf.p, qf.data(), qf.size() * sizeof(float))) return rc;
        ctx->qf_max = qmax;
        // rows the contraction reads, by its own test (incore.hip, store_skip)
        const double thr = ctx->opt.schwarz_threshold * (1.0 - 1e-6);
        int64_t live = 0;
        for (float f : qf) live += (double)f * qmax >= thr ? 1 : 0;
        ctx->qf_live = live;
        ctx->qf_valid = true;
    }
    // task sets of the previous molecule: marked stale (rebuilt on first use), device buffers kept
    auto stale = [](TaskSet& ts) {
        ts.threshold = -1;
        ts.classes.clear();
    };
    stale(ctx->tasks_all);
    stale(ctx->tasks_incore);
    for (auto& kv : ctx->tasks_by_level) stale(kv.second);
    for (auto& kv : ctx->tasks_mixed_hi) stale(kv.second);
    for (auto& kv : ctx->tasks_mixed_lo) stale(kv.second);
    ctx->oe_valid = false;
    ctx->have_mol = true;
    return HF_OK;
}

int ensure_tasks(hf_ctx* ctx, double thr, TaskSet& ts, double upper = 0.0, bool sliced = true) {
    if (ts.threshold == thr && ts.upper == upper && ts.sliced == sliced && !ts.classes.empty()) return HF_OK;
    return hf_build_tasks(ctx, thr, ts, upper, sliced);
}

// Threshold grid 10^(k/4): level of thr rounded down (never screens more than asked); thresholds
// <= 0 (no screening) map to kNoScreen.
constexpr int kNoScreen = -100000;
int threshold_level(double thr) {
    if (!(thr > 0)) return kNoScreen;
    return std::max(kNoScreen + 1, (int)std::floor(4.0 * std::log10(thr) + 1e-9));
}
double level_threshold(int level) { return level == kNoScreen ? 0.0 : std::pow(10.0, level / 4.0); }

// Screened task set for threshold thr, built once per level and molecule.
int screened_tasks(hf_ctx* ctx, double thr, TaskSet*& out) {
    const int level = threshold_level(thr);
    TaskSet& ts = ctx->tasks_by_level[level];
    out = &ts;
    return ensure_tasks(ctx, level_threshold(level), ts);
}

// Mixed precision: the tasks at threshold thr split at the FP32 bound t32 into the FP64 part
// (Q_p Q_q >= t32) and the FP32 band (thr <= Q_p Q_q < t32).
int mixed_tasks(hf_ctx* ctx, double thr, double t32, TaskSet*& hi, TaskSet*& lo) {
    const int level = threshold_level(thr);
    const double t = level_threshold(level);
    const int64_t key = (int64_t)level * 1000000 + threshold_level(t32);
    hi = &ctx->tasks_mixed_hi[key];
    lo = &ctx->tasks_mixed_lo[key];
    if (!(t32 > t)) {                     // nothing below the FP32 bound: all FP64
        if (int rc = ensure_tasks(ctx, t, *hi)) return rc;
        return ensure_tasks(ctx, std::numeric_limits<double>::max(), *lo);     // empty
    }
    if (int rc = ensure_tasks(ctx, t32, *hi)) return rc;
    return ensure_tasks(ctx, t, *lo, t32);
}

}  // namespace

// Expanded task list of one class: entry b * 256 + i is task i of launch block b, (-1, -1) in a
// segment's padding.
// With slices (tl.slice_prims > 0) a (bra, ket group) block of cnt kets with S slices holds cnt * S
// tasks, slice-major (the lanes of a wavefront share a slice and hold consecutive kets).
__global__ void __launch_bounds__(256) expand_tasks_kernel(TaskList tl, int2* __restrict__ out) {
    const TaskSeg sg = tl.seg[segment_of(tl.seg, tl.nseg, blockIdx.x, true)];
    const int64_t t = ((int64_t)blockIdx.x - sg.block0) * blockDim.x + threadIdx.x;
    int2 r = make_int2(-1, -1);
    if (t < sg.ntasks) {
        const int64_t* __restrict__ pre = tl.prefix + sg.prefix_off;
        const int pl = bra_of_task(pre, t, 0, sg.nbra);
        r.x = sg.bra0 + pl;
        if (tl.slice_prims <= 0) {
            r.y = ket_of_task(tl, sg, pre, pl, t);
        } else {
            const int* ko = tl.koff + sg.koff_off + (int64_t)pl * (sg.nkg + 1);
            const int* kb = tl.kbeg + sg.kbeg_off + (int64_t)pl * sg.nkg;
            const int rank = (int)(t - pre[pl]);
            int g = 0;
            while (g + 1 < sg.nkg && ko[g + 1] <= rank) ++g;
            const int npb = tl.bra.rec[r.x].nprim, npk = tl.ket.rec[tl.kgroup[g]].nprim;
            const int S = slice_count(npb, npk, tl.slice_prims);
            const int cnt = (ko[g + 1] - ko[g]) / S, lr = rank - ko[g];
            const int sl = lr / cnt, q = tl.kgroup[g] + kb[g] + (lr - sl * cnt);
            r.y = S > 1 ? q | ((sl + 1) << kTaskKetBits) : q;
        }
    }
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = r;
}

// J/K task sets slice heavily contracted quartets (kSliceTargetPrims primitive quartets per task)
// in classes with fewer than kSliceClassMaxQuartets quartets; larger classes fill the chip with
// whole quartets, and slicing them would only multiply their digestion atomics.
constexpr int64_t kSliceClassMaxQuartets = 1 << 18;
constexpr int kSliceTargetPrims = 16;

int hf_build_tasks(hf_ctx* ctx, double thr, TaskSet& ts, double upper, bool sliced) {
    ts.classes.clear();
    ts.sliced = sliced;
    std::vector<int64_t> prefix_all;
    std::vector<int> koff_all, kbeg_all;
    std::vector<TaskSeg> seg_all;
    // count of q in [k0, k1) (Q descending) with Q_q >= x
    auto head = [](const PairClassHost& K, int k0, int k1, double x) {
        return (int)(std::upper_bound(K.Q.begin() + k0, K.Q.begin() + k1, x, [](double v, double y) { return v > y; }) -
                     (K.Q.begin() + k0));
    };
    for (int c = 0; c < kNumClasses; ++c) {
        const ClassInfo& ci = kClasses[c];
        QuartetClassTasks qc;
        qc.cls = c;
        qc.bra = pair_class_index(ci.la, ci.lb);
        qc.ket = pair_class_index(ci.lc, ci.ld);
        const PairClassHost& B = ctx->pcls[qc.bra];
        const PairClassHost& K = ctx->pcls[qc.ket];
        const bool same = qc.bra == qc.ket;
        qc.seg_off = seg_all.size();
        const int nbg = B.group.empty() ? 0 : (int)B.group.size() - 1;
        const int nkg = K.group.empty() ? 0 : (int)K.group.size() - 1;
        // kets [beg, end) of ket group kg (ranks within the group) for bra p of bra group bg
        auto range = [&](int p, int bg, int kg, int& beg, int& end) {
            const int b0 = B.group[bg], k0 = K.group[kg], k1 = K.group[kg + 1];
            beg = end = 0;
            if (same && kg > bg) return;        // one class: each unordered pair of pairs once (q <= p)
            end = k1 - k0;
            if (thr > 0) end = B.Q[p] > 0 ? head(K, k0, k1, thr / B.Q[p]) : 0;
            if (upper > 0) beg = B.Q[p] > 0 ? head(K, k0, k1, upper / B.Q[p]) : 0;
            if (same && kg == bg) {
                end = std::min(end, p - b0 + 1);
                beg = std::min(beg, end);
            }
            if (beg > end) beg = end;
        };
        // primitive slices only where the class has too few quartets to fill the chip by itself
        if (sliced && ci.la + ci.lb + ci.lc + ci.ld < kCoopMinL) {
            int64_t nq = 0;
            for (int bg = 0; bg < nbg; ++bg)
                for (int p = B.group[bg]; p < B.group[bg + 1]; ++p)
                    for (int kg = 0; kg < nkg; ++kg) {
                        int beg, end;
                        range(p, bg, kg, beg, end);
                        nq += end - beg;
                    }
            qc.slice_prims = nq < kSliceClassMaxQuartets ? kSliceTargetPrims : 0;
        }
        for (int bg = 0; bg < nbg; ++bg) {
            const int b0 = B.group[bg], b1 = B.group[bg + 1];
            TaskSeg sg{};
            sg.bra0 = b0;
            sg.nbra = b1 - b0;
            sg.nkg = nkg;
            sg.prefix_off = (int64_t)prefix_all.size();
            sg.koff_off = (int64_t)koff_all.size();
            sg.kbeg_off = (int64_t)kbeg_all.size();
            int64_t acc = 0;
            prefix_all.push_back(0);
            for (int p = b0; p < b1; ++p) {
                int own = 0;
                koff_all.push_back(0);
                for (int kg = 0; kg < nkg; ++kg) {
                    const int k0 = K.group[kg];
                    int beg, end;
                    range(p, bg, kg, beg, end);
                    const int cnt = end - beg;
                    own += cnt * slice_count(B.nprim[p], K.nprim[k0], qc.slice_prims);
                    qc.quartets += cnt;
                    koff_all.push_back(own);
                    kbeg_all.push_back(beg);
                    qc.prim_quartets += (int64_t)cnt * B.nprim[p] * K.nprim[k0];
                }
                acc += own;
                prefix_all.push_back(acc);
            }
            if (acc == 0) {
                prefix_all.resize(sg.prefix_off);
                koff_all.resize(sg.koff_off);
                kbeg_all.resize(sg.kbeg_off);
                continue;
            }
            sg.ntasks = acc;
            sg.task0 = qc.ntasks;
            sg.block0 = qc.nblocks;
            qc.ntasks += acc;
            qc.nblocks += (acc + 255) / 256;
            seg_all.push_back(sg);
        }
        qc.nseg = (int)(seg_all.size() - qc.seg_off);
        ts.classes.push_back(qc);
    }
    HF_HIP_CHECK(ctx, ts.d_prefix.alloc(std::max<size_t>(prefix_all.size(), 1)));
    HF_HIP_CHECK(ctx, ts.d_koff.alloc(std::max<size_t>(koff_all.size(), 1)));
    HF_HIP_CHECK(ctx, ts.d_kbeg.alloc(std::max<size_t>(kbeg_all.size(), 1)));
    HF_HIP_CHECK(ctx, ts.d_seg.alloc(std::max<size_t>(seg_all.size(), 1)));
    if (!prefix_all.empty())
        if (int rc = hf_upload(ctx, ts.d_prefix.p, prefix_all.data(), prefix_all.size() * sizeof(int64_t))) return rc;
    if (!koff_all.empty())
        if (int rc = hf_upload(ctx, ts.d_koff.p, koff_all.data(), koff_all.size() * sizeof(int))) return rc;
    if (!kbeg_all.empty())
        if (int rc = hf_upload(ctx, ts.d_kbeg.p, kbeg_all.data(), kbeg_all.size() * sizeof(int))) return rc;
    if (!seg_all.empty())
        if (int rc = hf_upload(ctx, ts.d_seg.p, seg_all.data(), seg_all.size() * sizeof(TaskSeg))) return rc;
    // expand every class's tasks once (launch-order (bra, ket) pairs)
    size_t total = 0;
    for (QuartetClassTasks& qc : ts.classes) {
        qc.tasks_off = total;
        total += (size_t)qc.nblocks * 256;
    }
    HF_HIP_CHECK(ctx, ts.d_tasks.alloc(std::max<size_t>(total, 1)));
    for (const QuartetClassTasks& qc : ts.classes) {
        if (qc.nblocks == 0) continue;
        TaskList tl = task_list(ctx, ts, qc);
        hipLaunchKernelGGL(expand_tasks_kernel, dim3((unsigned)qc.nblocks), dim3(256), 0, ctx->stream, tl,
                           ts.d_tasks.p + qc.tasks_off);
    }
    HF_HIP_CHECK(ctx, hipGetLastError());
    HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ts.threshold = thr;
    ts.upper = upper;
    return HF_OK;
}

int hf_jk_device_impl(hf_ctx* ctx, int ndens, const double* d_D, double* d_J, double* d_K, hipStream_t st, double thr,
                      bool accumulate, bool fp32) {
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (ndens < 1 || ndens > 2) return ctx->fail(HF_E_ARG, "ndens must be 1 or 2");
    const int n = ctx->mol.nbf;
    const size_t nn = (size_t)n * n;
    if (thr <= 0) thr = ctx->opt.schwarz_threshold;
    std::vector<std::pair<int, const TaskSet*>> jobs;
    if (fp32) {
        TaskSet *hi = nullptr, *lo = nullptr;
        if (int rc = mixed_tasks(ctx, thr, ctx->opt.fp32_threshold, hi, lo)) return rc;
        jobs.push_back({MODE_JK, hi});
        jobs.push_back({MODE_JK32, lo});
    } else {
        TaskSet* ts = nullptr;
        if (int rc = screened_tasks(ctx, thr, ts)) return rc;
        jobs.push_back({MODE_JK, ts});
    }
    HF_HIP_CHECK(ctx, ctx->d_Jp.alloc(nn));
    HF_HIP_CHECK(ctx, ctx->d_Kp.alloc(2 * nn));
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(nn)), dim3(256), 0, st, ctx->d_Jp.p, nn, 0.0);
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(ndens * nn)), dim3(256), 0, st, ctx->d_Kp.p, ndens * nn, 0.0);
    SinkArgs sink{};
    sink.n = n;
    sink.ndens = ndens;
    sink.D = d_D;
    sink.J = ctx->d_Jp.p;
    sink.K = ctx->d_Kp.p;
    sink.boys32 = ctx->d_boys32.p;
    const size_t nfx = (1 + (size_t)ndens) * nn;   // J' then K' elements
    if (ctx->opt.deterministic) {
        HF_HIP_CHECK(ctx, ctx->d_fx.alloc(2 * 3 * nn));
        hipLaunchKernelGGL(fill_kernel, dim3(grid_for(2 * nfx)), dim3(256), 0, st, (double*)ctx->d_fx.p, 2 * nfx, 0.0);
        sink.fx = ctx->d_fx.p;
    }
    if (int rc = run_classes(ctx, jobs, sink, st, true)) return rc;
    if (ctx->opt.deterministic) {
        hipLaunchKernelGGL(fx_finalize_kernel, dim3(grid_for(nn)), dim3(256), 0, st, ctx->d_fx.p, ctx->d_Jp.p, nn);
        hipLaunchKernelGGL(fx_finalize_kernel, dim3(grid_for(ndens * nn)), dim3(256), 0, st, ctx->d_fx.p + 2 * nn,
                           ctx->d_Kp.p, ndens * nn);
    }
    hipLaunchKernelGGL(symmetrize_kernel, dim3(grid_for(nn)), dim3(256), 0, st, ctx->d_Jp.p, d_J, n, 1,
                       accumulate ? 1 : 0);
    hipLaunchKernelGGL(symmetrize_kernel, dim3(grid_for(ndens * nn)), dim3(256), 0, st, ctx->d_Kp.p, d_K, n, ndens,
                       accumulate ? 1 : 0);
    HF_HIP_CHECK(ctx, hipGetLastError());
    return HF_OK;
}

int hf_eri_pairs_impl(hf_ctx* ctx, double* d_G, uint64_t len, hipStream_t st) {
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    const uint64_t n = (uint64_t)ctx->mol.nbf, np = (uint64_t)pair_pad_base((int64_t)n);
    if (!d_G || len < np * np) return ctx->fail(HF_E_ARG, "pair-matrix buffer too small");
    // whole quartets (the store needs complete values), screened at the context's threshold
    if (int rc = ensure_tasks(ctx, ctx->opt.schwarz_threshold, ctx->tasks_incore, 0.0, false)) return rc;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(np * np)), dim3(256), 0, st, d_G, (size_t)(np * np), 0.0);
    SinkArgs sink{};
    sink.n = (int)n;
    sink.packed = d_G;
    sink.packed_len = np * np;
    sink.pair_stride = np;
    return run_classes(ctx, MODE_PACKED, ctx->tasks_incore, sink, st, true);
}

int hf_stream_enter(hf_ctx* ctx, hipStream_t st) {
    if (st == ctx->stream) return HF_OK;
    for (hipEvent_t& e : ctx->order_ev)
        if (!e) HF_HIP_CHECK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HF_HIP_CHECK(ctx, hipEventRecord(ctx->order_ev[0], ctx->stream));
    HF_HIP_CHECK(ctx, hipStreamWaitEvent(st, ctx->order_ev[0], 0));
    return HF_OK;
}

int hf_stream_leave(hf_ctx* ctx, hipStream_t st) {
    if (st == ctx->stream) return HF_OK;
    HF_HIP_CHECK(ctx, hipEventRecord(ctx->order_ev[1], st));
    HF_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, ctx->order_ev[1], 0));
    return HF_OK;
}

extern "C" {

void hf_default_options(hf_options* o) {
    o->device = 0;
    o->schwarz_threshold = 1e-12;
    o->prim_threshold = 1e-20;
    o->exact_transfer = 0;
    o->precision = HF_PREC_FP64;
    o->fp32_threshold = 1e-6;
    o->profile = 0;
    o->profile_class = -1;
    o->incore_shape = -1;
    o->deterministic = 0;
}

int hf_create(const hf_options* opt, hf_ctx** out) {
    if (!out) return HF_E_ARG;
    *out = nullptr;
    hf_ctx* ctx = new (std::nothrow) hf_ctx();
    if (!ctx) return HF_E_OOM;
    if (opt) ctx->opt = *opt; else hf_default_options(&ctx->opt);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete ctx;
        return HF_E_HIP;
    }
    if (ctx->opt.device < 0 || ctx->opt.device >= ndev || hipSetDevice(ctx->opt.device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return HF_E_HIP;
    }
    if (rocblas_create_handle(&ctx->blas) != rocblas_status_success) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return HF_E_HIP;
    }
    rocblas_set_stream(ctx->blas, ctx->stream);
    // the allocation-free entry point's scratch on this device (its address, taken once per device)
    if (hf_incore_scratch_init(ctx) != HF_OK) {
        rocblas_destroy_handle(ctx->blas);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return HF_E_HIP;
    }
    // pinned buffer for the SCF's small read-backs (without it they go through pageable memory)
    if (hipHostMalloc((void**)&ctx->hpin, kPinnedDoubles * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        ctx->hpin = nullptr;
        (void)hipGetLastError();
    }
    *out = ctx;
    return HF_OK;
}

int hf_destroy(hf_ctx* ctx) {
    if (!ctx) return HF_OK;
    (void)hipSetDevice(ctx->opt.device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->side_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->order_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->incore_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t s2 : ctx->side) (void)hipStreamDestroy(s2);
    if (ctx->eri_stream) {
        (void)hipStreamSynchronize(ctx->eri_stream);
        (void)hipStreamDestroy(ctx->eri_stream);
    }
    if (ctx->hpin) (void)hipHostFree(ctx->hpin);
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    hf_scf_session_free(ctx);
    hf_scf_workspace_free(ctx);
    if (ctx->blas) rocblas_destroy_handle(ctx->blas);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HF_OK;
}

const char* hf_last_error(const hf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hf_set_molecule(hf_ctx* ctx, const char* basis_file, int natoms, const int* Z, const double* xyz) {
    if (!ctx || !basis_file || natoms <= 0 || !Z || !xyz) return ctx ? ctx->fail(HF_E_ARG, "bad arguments") : HF_E_ARG;
    ctx->have_mol = false;
    for (int k = 0; k < 3 * natoms; ++k)
        if (!std::isfinite(xyz[k])) return ctx->fail(HF_E_ARG, "atom coordinate is not finite");
    std::string err;
    if (!hfhost::load_molecule(basis_file, natoms, Z, xyz, ctx->mol, err)) return ctx->fail(HF_E_IO, err);
    return setup_molecule(ctx);
}

int hf_set_basis(hf_ctx* ctx, const hf_basis_soa* b) {
    if (!ctx || !b) return ctx ? ctx->fail(HF_E_ARG, "null basis") : HF_E_ARG;
    ctx->have_mol = false;
    std::string err;
    if (!hfhost::molecule_from_soa(b, ctx->mol, err)) return ctx->fail(HF_E_ARG, err);
    return setup_molecule(ctx);
}

int hf_nbf(const hf_ctx* ctx) { return ctx && ctx->have_mol ? ctx->mol.nbf : 0; }
int hf_nshell(const hf_ctx* ctx) { return ctx && ctx->have_mol ? (int)ctx->mol.shells.size() : 0; }

int hf_electrons(const hf_ctx* ctx, int* alpha, int* beta) {
    if (!ctx || !ctx->have_mol) return HF_E_STATE;
    const int ne = ctx->mol.electrons();
    const bool set = ctx->mol.alpha_set > 0 || ctx->mol.beta_set > 0;
    if (alpha) *alpha = set ? ctx->mol.alpha_set : ne / 2;
    if (beta) *beta = set ? ctx->mol.beta_set : ne - ne / 2;
    return ne;
}

int hf_set_electrons(hf_ctx* ctx, int alpha, int beta) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (alpha < 0 || beta < 0) return ctx->fail(HF_E_ARG, "negative electron count");
    ctx->mol.alpha_set = alpha;
    ctx->mol.beta_set = beta;
    return HF_OK;
}

int hf_set_electric_field(hf_ctx* ctx, const double* field) {
    if (!ctx || !field) return ctx ? ctx->fail(HF_E_ARG, "null field") : HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(field[k])) return ctx->fail(HF_E_ARG, "field component is not finite");
        ctx->mol.field[k] = field[k];
    }
    return HF_OK;
}

double hf_nuclear_electric_field_energy(const hf_ctx* ctx) {
    // Molecule::NuclearElectricFieldEnergy (HartreeFock/Molecule.cpp:86-94): -sum_A Z_A R_A . E
    if (!ctx || !ctx->have_mol) return 0.0;
    double e = 0;
    for (size_t a = 0; a < ctx->mol.Z.size(); ++a)
        for (int k = 0; k < 3; ++k) e -= ctx->mol.Z[a] * ctx->mol.xyz[3 * a + k] * ctx->mol.field[k];
    return e;
}

double hf_nuclear_repulsion(const hf_ctx* ctx) {
    return ctx && ctx->have_mol ? hfhost::nuclear_repulsion(ctx->mol) : 0.0;
}

uint64_t hf_packed_index(uint32_t i, uint32_t j, uint32_t k, uint32_t l) {
    auto two = [](uint64_t a, uint64_t b) { return a < b ? b * (b + 1) / 2 + a : a * (a + 1) / 2 + b; };
    return two(two(i, j), two(k, l));
}

uint64_t hf_packed_len(int n) { return hf_packed_index(n, n, n, n) + 1; }

int hf_eri_packed_device(hf_ctx* ctx, double* d_out, uint64_t len, void* stream) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    const int n = ctx->mol.nbf;
    const uint64_t need = hf_packed_index(n - 1, n - 1, n - 1, n - 1) + 1;
    if (!d_out || len < need) return ctx->fail(HF_E_ARG, "packed buffer too small");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    int rc = ensure_tasks(ctx, 0.0, ctx->tasks_all, 0.0, false);   // packed store: whole quartets
    if (rc) return rc;
    if ((rc = hf_stream_enter(ctx, st))) return rc;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(len)), dim3(256), 0, st, d_out, (size_t)len, 0.0);
    SinkArgs sink{};
    sink.n = n;
    sink.packed = d_out;
    sink.packed_len = len;
    if ((rc = run_classes(ctx, MODE_PACKED, ctx->tasks_all, sink, st, true))) return rc;
    return hf_stream_leave(ctx, st);
}

int hf_eri_packed(hf_ctx* ctx, double* out, uint64_t len) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!out) return ctx->fail(HF_E_ARG, "null output");
    DevBuf<double> d;
    HF_HIP_CHECK(ctx, d.alloc(len));
    int rc = hf_eri_packed_device(ctx, d.p, len, nullptr);
    if (rc) return rc;
    HF_HIP_CHECK(ctx, hipMemcpyAsync(out, d.p, len * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return HF_OK;
}

uint64_t hf_pair_matrix_len(int nbf) {
    const uint64_t np = nbf > 0 ? (uint64_t)pair_pad_base(nbf) : 0;
    return np * np;
}

int hf_eri_pairs_device(hf_ctx* ctx, double* d_G, uint64_t len, void* stream) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    // the task list is built (and synchronised) on the context's stream before the caller's stream is ordered
    if (int rc = ensure_tasks(ctx, ctx->opt.schwarz_threshold, ctx->tasks_incore, 0.0, false)) return rc;
    if (int rc = hf_stream_enter(ctx, st)) return rc;
    if (int rc = hf_eri_pairs_impl(ctx, d_G, len, st)) return rc;
    return hf_stream_leave(ctx, st);
}

int hf_build_jk_incore_device(hf_ctx* ctx, int ndens, const double* d_G, const double* d_D, double* d_J, double* d_K,
                              void* stream) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!d_G || !d_D || !d_J || !d_K) return ctx->fail(HF_E_ARG, "null device buffer");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    if (int rc = hf_stream_enter(ctx, st)) return rc;
    if (int rc = hf_incore_jk_impl(ctx, ndens, d_G, nullptr, d_D, d_J, d_K, st, false, false)) return rc;
    return hf_stream_leave(ctx, st);
}

int hf_build_jk_incore(hf_ctx* ctx, int ndens, const double* D, double* J, double* K) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!D || !J || !K || ndens < 1 || ndens > 2) return ctx->fail(HF_E_ARG, "bad arguments");
    const int n = ctx->mol.nbf;
    if (n > kIncoreMaxN) return ctx->fail(HF_E_UNSUPPORTED, "in-core J/K takes at most 512 basis functions");
    const size_t nn = (size_t)n * n;
    const uint64_t len = hf_pair_matrix_len(n);
    HF_HIP_CHECK(ctx, ctx->d_G.alloc(len));
    if (int rc = hf_eri_pairs_impl(ctx, ctx->d_G.p, len, ctx->stream)) return rc;
    HF_HIP_CHECK(ctx, ctx->d_D.alloc(2 * nn));
    HF_HIP_CHECK(ctx, ctx->d_J.alloc(nn));
    HF_HIP_CHECK(ctx, ctx->d_K.alloc(2 * nn));
    HF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_D.p, D, ndens * nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = hf_incore_jk_impl(ctx, ndens, ctx->d_G.p, nullptr, ctx->d_D.p, ctx->d_J.p, ctx->d_K.p, ctx->stream, false,
                               true);
    if (rc) return rc;
    HF_HIP_CHECK(ctx, hipMemcpyAsync(J, ctx->d_J.p, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HF_HIP_CHECK(ctx, hipMemcpyAsync(K, ctx->d_K.p, ndens * nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return HF_OK;
}

int hf_build_jk_device(hf_ctx* ctx, int ndens, const double* d_D, double* d_J, double* d_K, void* stream) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!d_D || !d_J || !d_K) return ctx->fail(HF_E_ARG, "null device buffer");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    // task lists are built (and synchronised) on the context's stream before the caller's stream is ordered
    TaskSet *ts = nullptr, *lo = nullptr;
    if (ctx->opt.precision == HF_PREC_MIXED) {
        if (int rc = mixed_tasks(ctx, ctx->opt.schwarz_threshold, ctx->opt.fp32_threshold, ts, lo)) return rc;
    } else if (int rc = screened_tasks(ctx, ctx->opt.schwarz_threshold, ts)) {
        return rc;
    }
    if (int rc = hf_stream_enter(ctx, st)) return rc;
    if (int rc = hf_jk_device_impl(ctx, ndens, d_D, d_J, d_K, st, 0.0, false, ctx->opt.precision == HF_PREC_MIXED))
        return rc;
    return hf_stream_leave(ctx, st);
}

int hf_build_jk(hf_ctx* ctx, int ndens, const double* D, double* J, double* K) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!D || !J || !K || ndens < 1 || ndens > 2) return ctx->fail(HF_E_ARG, "bad arguments");
    const size_t nn = (size_t)ctx->mol.nbf * ctx->mol.nbf;
    HF_HIP_CHECK(ctx, ctx->d_D.alloc(2 * nn));
    HF_HIP_CHECK(ctx, ctx->d_J.alloc(nn));
    HF_HIP_CHECK(ctx, ctx->d_K.alloc(2 * nn));
    HF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_D.p, D, ndens * nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = hf_jk_device_impl(ctx, ndens, ctx->d_D.p, ctx->d_J.p, ctx->d_K.p, ctx->stream, 0.0, false,
                               ctx->opt.precision == HF_PREC_MIXED);
    if (rc) return rc;
    HF_HIP_CHECK(ctx, hipMemcpyAsync(J, ctx->d_J.p, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HF_HIP_CHECK(ctx, hipMemcpyAsync(K, ctx->d_K.p, ndens * nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return HF_OK;
}

int hf_one_electron(hf_ctx* ctx, double* S, double* T, double* V) {
    if (!ctx) return HF_E_ARG;
    if (!ctx->have_mol) return ctx->fail(HF_E_STATE, "no molecule set");
    if (!S || !T || !V) return ctx->fail(HF_E_ARG, "null output");
    const size_t nn = (size_t)ctx->mol.nbf * ctx->mol.nbf;
    DevBuf<double> d;
    HF_HIP_CHECK(ctx, d.alloc(3 * nn));
    if (int rc = hf_one_electron_device(ctx, d.p, d.p + nn, d.p + 2 * nn, ctx->stream)) return rc;
    HF_HIP_CHECK(ctx, hipMemcpyAsync(S, d.p, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HF_HIP_CHECK(ctx, hipMemcpyAsync(T, d.p + nn, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HF_HIP_CHECK(ctx, hipMemcpyAsync(V, d.p + 2 * nn, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return HF_OK;
}

int hf_kernel_stats(const hf_ctx* ctx, hf_kernel_stat* out, int max) {
    if (!ctx || !out) return 0;
    if (flush_kernel_times(const_cast<hf_ctx*>(ctx)) != HF_OK) return 0;
    int k = 0;
    for (const KernelStat& s : ctx->stats) {
        if (k >= max) break;
        if (s.launches == 0) continue;
        const ClassInfo& ci = kClasses[s.cls];
        out[k].la = ci.la;
        out[k].lb = ci.lb;
        out[k].lc = ci.lc;
        out[k].ld = ci.ld;
        out[k].launches = s.launches;
        out[k].quartets = s.quartets;
        out[k].prim_quartets = s.prim_quartets;
        out[k].eris = s.eris;
        out[k].flops = s.flops;
        out[k].ms = s.ms;
        out[k].fp32 = s.fp32;
        ++k;
    }
    // the in-core contraction, as class (-1 -1 | -1 -1)
    for (int f = 0; f < 2 && k < max; ++f) {
        const KernelStat& s = ctx->incore_stat[f];
        if (s.launches == 0) continue;
        out[k].la = out[k].lb = out[k].lc = out[k].ld = -1;
        out[k].launches = s.launches;
        out[k].quartets = 0;
        out[k].prim_quartets = 0;
        out[k].eris = s.eris;
        out[k].flops = s.flops;
        out[k].ms = s.ms;
        out[k].fp32 = f;
        ++k;
    }
    return k;
}

int hf_kernel_stats_reset(hf_ctx* ctx) {
    if (!ctx) return HF_E_ARG;
    if (int rc = flush_kernel_times(ctx)) return rc;
    for (KernelStat& s : ctx->stats) {
        const int c = s.cls, f = s.fp32;
        s = KernelStat{};
        s.cls = c;
        s.fp32 = f;
    }
    ctx->incore_stat[0] = ctx->incore_stat[1] = KernelStat{};
    return HF_OK;
}

}  // extern "C"
