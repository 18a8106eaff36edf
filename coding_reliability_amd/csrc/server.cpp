// Continuous-batching request loop with device-resident histogram state.
//
// One scheduler thread per server.  Each iteration:
//   1. admits queued requests (FIFO) while a slot is free and the summed
//      device footprint stays within the HBM budget; an admitted request gets
//      a histogram slot in the server's device slab, zeroed on the device;
//   2. gives every active request a slice of its remaining work items, sized
//      from a measured throughput model so the iteration lasts about
//      iteration_ms of device time, shared equally among active requests,
//      but never fewer than two items per wave of the launch (a launch with
//      fewer items than waves runs as long as one item however few there
//      are, so a smaller slice would only lower the measured rate and shrink
//      the next slice further);
//   3. launches one pair-kernel pass over the AN slices and one Hamming pass
//      over the Monte Carlo slices (each kernel works on all slices at once);
//   4. retires finished requests: copies their histogram to the host and
//      frees the slot.
// Requests arriving during an iteration are admitted at the next one.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cr_internal.h"

using namespace cr;

namespace {

constexpr uint32_t kSlotBins = 130;  // histogram slot size (uint64 counters)

// Clears the listed histogram slots (one launch for all slots freed since the
// last iteration; a slot is clean before its next request's first pass).
__global__ void zero_slots_kernel(unsigned long long *__restrict__ slab, const uint32_t *__restrict__ slots,
                                  uint32_t n) {
    for (uint32_t b = blockIdx.x; b < n; b += gridDim.x)
        for (uint32_t i = threadIdx.x; i < kSlotBins; i += blockDim.x) slab[(size_t)slots[b] * kSlotBins + i] = 0;
}

struct Request {
    uint64_t id = 0;
    cr_job job{};
    uint32_t kind = 0;
    cr_status status = CR_OK;
    int state = 0;  // 0 queued, 1 active, 2 done
    bool cancel = false;
    int slot = -1;
    uint64_t footprint = 0;
    uint32_t nbins = 0;
    uint64_t total_items = 0, cursor = 0;
    long double units_per_item = 0;
    uint64_t units = 0;
    PairJob pj{};
    HamJob hj{};
    // exhaustive requests cheaper through the code's spectrum: items are
    // batches of transforms, run after the pair and Hamming launches
    bool spectral = false;
    SpectralPlan sp{};
    DeviceBuffer sx, sslots;
    uint32_t iterations = 0;
    double t_submit = 0, t_admit = 0, t_finish = 0;
    std::vector<uint64_t> result;
};

}  // namespace

struct cr_server {
    cr_context *ctx = nullptr;
    cr_server_config cfg{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, evm = nullptr, ev1 = nullptr, ev2 = nullptr;
    DeviceBuffer slab, pair_jobs, ham_jobs, counter, slot_list;
    HostBuffer staging, readback;
    std::vector<int> free_slots;
    std::vector<uint32_t> dirty_slots;  // freed since the last iteration, cleared by it
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<uint64_t> active;
    uint64_t next_id = 1;
    uint64_t footprint = 0;
    cr_server_stats st{};
    bool stop = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;
    long double pair_rate = 8e9L;    // pairs per device ms (refined online)
    long double sample_rate = 1e9L;  // Hamming samples per device ms

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void loop();
    cr_status iterate(std::vector<Request *> &run);
};

namespace {

// Validates a job and fills its work description.
cr_status describe(cr_server *s, const cr_job &j, Request &r) {
    r.job = j;
    r.kind = j.kind;
    const uint32_t R = s->cfg.rows_per_lane;
    if (j.kind == CR_JOB_AN_EXHAUSTIVE || j.kind == CR_JOB_AN_GRID) {
        if (j.k < 1 || j.k > 40) return fail(CR_ERR_UNSUPPORTED, "AN request k=%u outside [1, 40]", j.k);
        if (j.A == 0) return fail(CR_ERR_INVALID_ARGUMENT, "A must be >= 1");
        uint64_t A = j.A, N = 1ull << j.k;
        if (j.kind == CR_JOB_AN_GRID) {
            if (j.stride == 0) return fail(CR_ERR_INVALID_ARGUMENT, "grid stride must be >= 1");
            if ((unsigned __int128)A * j.stride >> 64) return fail(CR_ERR_UNSUPPORTED, "A*stride exceeds 64 bits");
            A = j.A * j.stride;
            N = ((1ull << j.k) + j.stride - 1) / j.stride;
        }
        if ((unsigned __int128)(N - 1) * A >> 64) return fail(CR_ERR_UNSUPPORTED, "codewords exceed 64 bits");
        PairJob &p = r.pj;
        p = PairJob{};
        p.A = A;
        p.N = N;
        p.mode = pair_mode(A, N);
        p.nbins = pair_max_distance(A, N) + 1;
        // items of about 1/64 of the request (at most 2^24 pairs, a few ms for
        // one wave; at least one diagonal block): a small request spreads over
        // many waves instead of waiting on a few long items
        const uint64_t RB = 64ull * R;
        const long double pairs = (long double)N * (long double)(N - 1) / 2;
        const long double per_item = std::min<long double>(std::max<long double>(pairs / 64, (long double)RB * RB),
                                                            (long double)(1ull << 24));
        uint64_t Q = std::max<uint64_t>((uint64_t)(per_item / ((long double)RB * RB)), 1);
        pair_job_geometry(p, R, (uint32_t)Q);
        r.nbins = p.nbins;
        r.total_items = p.items;
        r.units = N * (N - 1) / 2;
        // the spectral path when it is cheaper and its buffers fit the budget
        // (otherwise the pair kernel, whose footprint is one histogram slot)
        uint32_t k = 0;
        if (j.kind == CR_JOB_AN_EXHAUSTIVE && spectral_preferred(A, N, &k) && spectral_plan(A, k, s->ctx->cus, &r.sp) &&
            r.sp.buffer_bytes + r.sp.slot_entries * sizeof(uint64_t) + 4096 <= s->cfg.hbm_budget_bytes) {
            r.spectral = true;
            r.total_items = r.sp.batches;
        }
    } else if (j.kind == CR_JOB_HAMMING_MC) {
        if (j.k < 1 || j.k > 64) return fail(CR_ERR_INVALID_ARGUMENT, "Hamming request k=%u outside [1, 64]", j.k);
        HamJob &h = r.hj;
        h = HamJob{};
        const uint32_t rbits = hamming_masks(j.k, h.masks);
        h.source = kHamMc;
        h.k = j.k;
        h.first = 0;
        h.count = j.samples;
        h.seed = j.seed;
        h.stream = j.stream;
        h.nbins = j.k + rbits + 1;
        h.lane_samples = kHamMaxLaneSamples;
        r.nbins = h.nbins;
        r.total_items = hamming_items(j.samples, h.lane_samples);
        r.units = j.samples;
    } else {
        return fail(CR_ERR_INVALID_ARGUMENT, "unknown request kind %u", j.kind);
    }
    r.units_per_item = r.total_items ? (long double)r.units / r.total_items : 1;
    r.footprint = ((uint64_t)kSlotBins * sizeof(uint64_t) + 255) / 256 * 256;
    if (r.spectral) r.footprint += r.sp.buffer_bytes + r.sp.slot_entries * sizeof(uint64_t);
    return CR_OK;
}

}  // namespace

cr_status cr_server::iterate(std::vector<Request *> &run) {
    // slices: equal share of the iteration's device time per request
    const long double budget_ms = cfg.iteration_ms;
    const long double share = budget_ms / std::max<size_t>(run.size(), 1);
    std::vector<PairJob> pjobs;
    std::vector<HamJob> hjobs;
    std::vector<std::pair<Request *, uint64_t>> took;
    uint64_t ptotal = 0, htotal = 0;
    uint32_t lds_pair = 2, lds_ham = 2;
    uint64_t n_an = 0, n_mc = 0;
    for (Request *r : run) {
        if (r->kind == CR_JOB_HAMMING_MC) {
            lds_ham = std::max(lds_ham, r->hj.nbins + 1);
            ++n_mc;
        } else if (!r->spectral) {
            lds_pair = std::max(lds_pair, r->pj.nbins + 1);
            ++n_an;
        }
    }
    std::vector<std::pair<Request *, uint64_t>> spec;
    // waves each launch runs with: a slice floor of two items per wave
    const uint64_t wp = n_an ? (uint64_t)pair_waves_per_cu(cfg.rows_per_lane, lds_pair) * ctx->cus : 0;
    const uint64_t wh = n_mc ? (uint64_t)hamming_waves_per_cu(lds_ham) * ctx->cus : 0;
    long double pair_units = 0, ham_units = 0;
    for (Request *r : run) {
        if (r->spectral) {
            // batches for the share at the plan's measured time per transformed element
            const long double batch_ms =
                (long double)r->sp.per_batch * (long double)(1ull << r->sp.k) * (long double)r->sp.sec_per_element * 1e3L;
            uint64_t n = std::max<uint64_t>((uint64_t)(share / batch_ms), 1);
            n = std::min<uint64_t>(n, r->total_items - r->cursor);
            spec.push_back({r, n});
            took.push_back({r, n});
            continue;
        }
        const bool mc = r->kind == CR_JOB_HAMMING_MC;
        const long double rate = mc ? sample_rate : pair_rate;
        uint64_t n = (uint64_t)(share * rate / std::max<long double>(r->units_per_item, 1));
        const uint64_t floor_n = mc ? (2 * wh + n_mc - 1) / n_mc : (2 * wp + n_an - 1) / n_an;
        n = std::max<uint64_t>({n, floor_n, 1});
        n = std::min<uint64_t>(n, r->total_items - r->cursor);
        unsigned long long *out = slab.as<unsigned long long>() + (size_t)r->slot * kSlotBins;
        if (r->kind == CR_JOB_HAMMING_MC) {
            HamJob h = r->hj;
            h.out = out;
            h.item_begin = htotal;
            h.items = n;
            h.item_offset = r->cursor;
            htotal += n;
            hjobs.push_back(h);
            ham_units += n * r->units_per_item;
        } else {
            PairJob p = r->pj;
            p.out = out;
            p.item_begin = ptotal;
            p.items = n;
            p.item_offset = r->cursor;
            ptotal += n;
            pjobs.push_back(p);
            pair_units += n * r->units_per_item;
        }
        took.push_back({r, n});
    }
    // a Hamming launch handles either 32-bit or 64-bit data words
    std::vector<HamJob> h32, h64;
    for (auto &h : hjobs) (h.k > 32 ? h64 : h32).push_back(h);
    auto rebase = [](std::vector<HamJob> &v) {
        uint64_t t = 0;
        for (auto &h : v) {
            h.item_begin = t;
            t += h.items;
        }
        return t;
    };
    const uint64_t t32 = rebase(h32), t64 = rebase(h64);
    std::vector<uint32_t> zero;
    {
        std::lock_guard<std::mutex> lk(mu);
        zero.swap(dirty_slots);
    }
    const size_t jbytes = pjobs.size() * sizeof(PairJob) + (h32.size() + h64.size()) * sizeof(HamJob);
    const size_t bytes = jbytes + zero.size() * sizeof(uint32_t);
    CR_TRY(staging.reserve(bytes));
    CR_TRY(pair_jobs.reserve(std::max<size_t>(pjobs.size(), 1) * sizeof(PairJob)));
    CR_TRY(ham_jobs.reserve(std::max<size_t>(h32.size() + h64.size(), 1) * sizeof(HamJob)));
    char *st8 = staging.as<char>();
    memcpy(st8, pjobs.data(), pjobs.size() * sizeof(PairJob));
    memcpy(st8 + pjobs.size() * sizeof(PairJob), h32.data(), h32.size() * sizeof(HamJob));
    memcpy(st8 + pjobs.size() * sizeof(PairJob) + h32.size() * sizeof(HamJob), h64.data(), h64.size() * sizeof(HamJob));
    memcpy(st8 + jbytes, zero.data(), zero.size() * sizeof(uint32_t));
    if (!zero.empty()) {
        CR_TRY(slot_list.reserve(zero.size() * sizeof(uint32_t)));
        CR_HIP(hipMemcpyAsync(slot_list.ptr, st8 + jbytes, zero.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              stream));
        const uint32_t blocks = (uint32_t)std::min<size_t>(zero.size(), 1024);
        hipLaunchKernelGGL(zero_slots_kernel, dim3(blocks), dim3(128), 0, stream, slab.as<unsigned long long>(),
                           slot_list.as<uint32_t>(), (uint32_t)zero.size());
        CR_HIP(hipGetLastError());
    }
    if (!pjobs.empty())
        CR_HIP(hipMemcpyAsync(pair_jobs.ptr, st8, pjobs.size() * sizeof(PairJob), hipMemcpyHostToDevice, stream));
    if (!h32.empty() || !h64.empty())
        CR_HIP(hipMemcpyAsync(ham_jobs.ptr, st8 + pjobs.size() * sizeof(PairJob),
                              (h32.size() + h64.size()) * sizeof(HamJob), hipMemcpyHostToDevice, stream));
    CR_HIP(hipEventRecord(ev0, stream));
    unsigned long long *ctr = counter.as<unsigned long long>();
    if (!pjobs.empty()) {
        PairLaunchConfig pc;
        pc.R = cfg.rows_per_lane;
        pc.lds_rows = lds_pair;
        CR_TRY(launch_pairs(pair_jobs.as<PairJob>(), (uint32_t)pjobs.size(), ptotal, ctr, pc, ctx->cus, stream,
                            nullptr, nullptr));
    }
    CR_HIP(hipEventRecord(evm, stream));
    if (!h32.empty())
        CR_TRY(launch_hamming(ham_jobs.as<HamJob>(), (uint32_t)h32.size(), t32, false, ctr + 1, false, lds_ham,
                              ctx->cus, stream, nullptr, nullptr));
    if (!h64.empty())
        CR_TRY(launch_hamming(ham_jobs.as<HamJob>() + h32.size(), (uint32_t)h64.size(), t64, true, ctr + 2, false,
                              lds_ham, ctx->cus, stream, nullptr, nullptr));
    CR_HIP(hipEventRecord(ev1, stream));
    for (auto &sr : spec) {
        Request *r = sr.first;
        if (!r->sx.ptr) {  // first slice: transform buffer and zeroed slots
            CR_TRY(r->sx.reserve(r->sp.buffer_bytes));
            CR_TRY(r->sslots.reserve(r->sp.slot_entries * sizeof(uint64_t)));
            CR_TRY(fill_zero_u64(r->sslots.as<unsigned long long>(), r->sp.slot_entries, stream));
        }
        CR_TRY(spectral_run(r->sp, r->sx.ptr, r->sslots.as<unsigned long long>(), r->cursor, sr.second, stream));
    }
    CR_HIP(hipEventRecord(ev2, stream));
    CR_HIP(hipStreamSynchronize(stream));
    float ms = 0, pms = 0, hms = 0;
    CR_HIP(hipEventElapsedTime(&ms, ev0, ev2));
    CR_HIP(hipEventElapsedTime(&pms, ev0, evm));
    CR_HIP(hipEventElapsedTime(&hms, evm, ev1));
    // refine the throughput model (exponential average) from launches that
    // had at least one item per wave, i.e. ran at the device's full width
    if (!pjobs.empty() && ptotal >= wp && pms > 0.05f)
        pair_rate = 0.7L * pair_rate + 0.3L * (pair_units / pms);
    if (!hjobs.empty() && htotal >= wh && hms > 0.05f)
        sample_rate = 0.7L * sample_rate + 0.3L * (ham_units / hms);
    std::lock_guard<std::mutex> lk(mu);
    st.iterations++;
    st.busy_device_ms += ms;
    for (auto &tk : took) {
        tk.first->cursor += tk.second;
        tk.first->iterations++;
    }
    return CR_OK;
}

void cr_server::loop() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::vector<Request *> run;
        {
            std::unique_lock<std::mutex> lk(mu);
            // retire cancelled active requests
            for (auto it = active.begin(); it != active.end();) {
                Request *r = reqs[*it].get();
                if (r->cancel) {
                    r->state = 2;
                    r->status = CR_ERR_CANCELLED;
                    r->t_finish = now();
                    r->sx.release();
                    r->sslots.release();
                    footprint -= r->footprint;
                    free_slots.push_back(r->slot);
                    dirty_slots.push_back((uint32_t)r->slot);
                    st.cancelled++;
                    it = active.erase(it);
                } else {
                    ++it;
                }
            }
            // admission: FIFO while a slot is free and the footprint fits
            while (!queue.empty()) {
                Request *r = reqs[queue.front()].get();
                if (r->cancel) {
                    r->state = 2;
                    r->status = CR_ERR_CANCELLED;
                    r->t_finish = now();
                    st.cancelled++;
                    queue.pop_front();
                    continue;
                }
                if (free_slots.empty() || footprint + r->footprint > cfg.hbm_budget_bytes) break;
                queue.pop_front();
                r->slot = free_slots.back();
                free_slots.pop_back();
                footprint += r->footprint;
                r->state = 1;
                r->t_admit = now();
                active.push_back(r->id);  // its slot is clean: zeroed at creation or when last freed
            }
            cv_done.notify_all();
            if (active.empty()) {
                if (stop) return;
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
                continue;
            }
            for (uint64_t id : active) run.push_back(reqs[id].get());
        }
        cr_status s = iterate(run);
        std::vector<Request *> fin;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto it = active.begin(); it != active.end();) {
                Request *r = reqs[*it].get();
                if (s != CR_OK || r->cursor >= r->total_items) {
                    fin.push_back(r);
                    it = active.erase(it);
                } else {
                    ++it;
                }
            }
        }
        // one copy of the slab range holding every finished histogram
        int lo = 0, hi = -1;
        if (!fin.empty()) {
            lo = hi = fin[0]->slot;
            for (Request *r : fin) {
                lo = std::min(lo, r->slot);
                hi = std::max(hi, r->slot);
            }
        }
        if (s == CR_OK && !fin.empty()) {
            hipError_t e = hipMemcpyAsync(readback.ptr, slab.as<unsigned long long>() + (size_t)lo * kSlotBins,
                                          (size_t)(hi - lo + 1) * kSlotBins * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                          stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) s = fail(CR_ERR_DEVICE, "result copy failed: %s", hipGetErrorString(e));
        }
        for (Request *r : fin) {
            std::vector<uint64_t> h(r->nbins, 0);
            cr_status rs = s;
            if (rs == CR_OK && r->spectral) {
                // the histogram from the request's slots (Krawtchouk transform on the host)
                std::vector<uint64_t> sl(r->sp.slot_entries);
                hipError_t e = hipMemcpy(sl.data(), r->sslots.ptr, sl.size() * sizeof(uint64_t), hipMemcpyDeviceToHost);
                rs = e == hipSuccess ? spectral_finish(r->sp, sl.data(), h.data(), r->nbins)
                                     : fail(CR_ERR_DEVICE, "spectral result copy failed: %s", hipGetErrorString(e));
            } else if (rs == CR_OK) {
                memcpy(h.data(), readback.as<uint64_t>() + (size_t)(r->slot - lo) * kSlotBins,
                       r->nbins * sizeof(uint64_t));
            }
            r->sx.release();
            r->sslots.release();
            std::lock_guard<std::mutex> lk(mu);
            r->result.swap(h);
            r->status = rs;
            r->state = 2;
            r->t_finish = now();
            footprint -= r->footprint;
            free_slots.push_back(r->slot);
            dirty_slots.push_back((uint32_t)r->slot);
            st.completed++;
        }
        cv_done.notify_all();
    }
}

extern "C" {

cr_status cr_server_create(cr_context *ctx, const cr_server_config *cfg, cr_server **out) {
    clear_error();
    if (!ctx || !out) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    CR_HIP(hipSetDevice(ctx->device));
    auto s = std::make_unique<cr_server>();
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.hbm_budget_bytes == 0) s->cfg.hbm_budget_bytes = 1ull << 30;
    if (s->cfg.max_active == 0) s->cfg.max_active = 64;
    if (s->cfg.rows_per_lane == 0) s->cfg.rows_per_lane = 8;
    if (s->cfg.iteration_ms <= 0) s->cfg.iteration_ms = 50;
    if (pair_waves_per_cu(s->cfg.rows_per_lane, 66) == 0)
        return fail(CR_ERR_INVALID_ARGUMENT, "rows_per_lane must be 2, 4, 8 or 16");
    // from here on a failure releases what was created (no worker yet)
    cr_server *raw = s.release();
    cr_status st = [raw]() -> cr_status {
        CR_HIP(hipStreamCreateWithFlags(&raw->stream, hipStreamNonBlocking));
        CR_HIP(hipEventCreate(&raw->ev0));
        CR_HIP(hipEventCreate(&raw->evm));
        CR_HIP(hipEventCreate(&raw->ev1));
        CR_HIP(hipEventCreate(&raw->ev2));
        const size_t slots = (size_t)raw->cfg.max_active * kSlotBins;
        CR_TRY(raw->slab.reserve(slots * sizeof(uint64_t)));
        CR_TRY(fill_zero_u64(raw->slab.as<unsigned long long>(), slots, raw->stream));
        CR_HIP(hipStreamSynchronize(raw->stream));
        CR_TRY(raw->counter.reserve(4 * sizeof(uint64_t)));
        CR_TRY(raw->readback.reserve(slots * sizeof(uint64_t)));
        return CR_OK;
    }();
    if (st != CR_OK) {
        const std::string msg = cr_last_error();
        cr_server_destroy(raw);
        return fail(st, "%s", msg.c_str());
    }
    for (int i = (int)raw->cfg.max_active - 1; i >= 0; --i) raw->free_slots.push_back(i);
    raw->t0 = std::chrono::steady_clock::now();
    raw->worker = std::thread([raw] { raw->loop(); });
    *out = raw;
    return CR_OK;
}

cr_status cr_server_submit(cr_server *s, const cr_job *job, uint64_t *id) {
    clear_error();
    if (!s || !job || !id) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    auto r = std::make_unique<Request>();
    CR_TRY(describe(s, *job, *r));
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->footprint > s->cfg.hbm_budget_bytes) {
        s->st.rejected++;
        return fail(CR_ERR_ADMISSION, "request needs %llu device bytes, budget is %llu",
                    (unsigned long long)r->footprint, (unsigned long long)s->cfg.hbm_budget_bytes);
    }
    if (s->stop) return fail(CR_ERR_CANCELLED, "server is shutting down");
    r->id = s->next_id++;
    r->t_submit = s->now();
    *id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->st.submitted++;
    s->cv_work.notify_one();
    return CR_OK;
}

cr_status cr_server_result(cr_server *s, uint64_t id, uint64_t *hist, uint32_t nbins, cr_job_result *res) {
    clear_error();
    if (!s) return fail(CR_ERR_INVALID_ARGUMENT, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(CR_ERR_NOT_FOUND, "unknown request %llu", (unsigned long long)id);
    Request *r = it->second.get();
    if (res) {
        memset(res, 0, sizeof(*res));
        res->done = r->state == 2;
        res->status = r->status;
        res->nbins = r->nbins;
        res->iterations = r->iterations;
        res->submit_s = r->t_submit;
        res->admit_s = r->t_admit;
        res->finish_s = r->t_finish;
        res->units = r->units;
    }
    if (r->state != 2) return fail(CR_ERR_BUSY, "request %llu not finished", (unsigned long long)id);
    if (r->status != CR_OK) return fail((cr_status)r->status, "request %llu failed", (unsigned long long)id);
    if (hist) {
        if (nbins < r->nbins) return fail(CR_ERR_INVALID_ARGUMENT, "nbins=%u < %u", nbins, r->nbins);
        memset(hist, 0, nbins * sizeof(uint64_t));
        memcpy(hist, r->result.data(), r->nbins * sizeof(uint64_t));
    }
    return CR_OK;
}

cr_status cr_server_wait(cr_server *s, uint64_t id, double timeout_ms) {
    clear_error();
    if (!s) return fail(CR_ERR_INVALID_ARGUMENT, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(CR_ERR_NOT_FOUND, "unknown request %llu", (unsigned long long)id);
    Request *r = it->second.get();
    auto pred = [&] { return r->state == 2; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, pred);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred)) {
        return fail(CR_ERR_BUSY, "request %llu still running", (unsigned long long)id);
    }
    return CR_OK;
}

cr_status cr_server_cancel(cr_server *s, uint64_t id) {
    clear_error();
    if (!s) return fail(CR_ERR_INVALID_ARGUMENT, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(CR_ERR_NOT_FOUND, "unknown request %llu", (unsigned long long)id);
    if (it->second->state != 2) it->second->cancel = true;
    s->cv_work.notify_one();
    return CR_OK;
}

cr_status cr_server_stats_get(cr_server *s, cr_server_stats *out) {
    clear_error();
    if (!s || !out) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->active = s->active.size();
    out->queued = s->queue.size();
    out->footprint_bytes = s->footprint;
    return CR_OK;
}

double cr_server_now(cr_server *s) { return s ? s->now() : 0.0; }

cr_status cr_server_destroy(cr_server *s) {
    if (!s) return CR_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto &kv : s->reqs)
            if (kv.second->state != 2) kv.second->cancel = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto &kv : s->reqs) {
        kv.second->sx.release();
        kv.second->sslots.release();
    }
    s->slab.release();
    s->pair_jobs.release();
    s->ham_jobs.release();
    s->counter.release();
    s->slot_list.release();
    s->staging.release();
    s->readback.release();
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->evm) (void)hipEventDestroy(s->evm);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->ev2) (void)hipEventDestroy(s->ev2);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return CR_OK;
}

}  // extern "C"
