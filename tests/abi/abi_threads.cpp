// abi_threads.cpp -- a C++ caller of the lmx C ABI under the reference's
// threading: the scalar Accel3::Intersect called once per ray from many
// worker threads at once (src/liblightmetrica/scheduler.cpp:146-183 hands
// grains of 10000 samples to TBB workers, each of which calls
// Scene3::Intersect -> Accel3::Intersect per ray, scene3.cpp:458-483).
//
//   abi_threads <mesh.bin> <rays.bin> <hits_out.bin> <threads> <isect>
//
// mesh.bin: uint32 nvert, uint32 ntri, nvert x 3 float32, ntri x 3 uint32.
// rays.bin: uint32 n, n x 8 float32 (lmx_ray).  hits_out.bin: n x 16 bytes
// (lmx_hit) from the per-ray, multi-threaded calls.
//
// Prints one JSON line: batched and per-call nanoseconds per ray on one
// thread, the threaded run's wall time, and the number of threaded hits that
// differ bit for bit from one batched single-thread call.  Exit status 0 when
// there is no difference.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "lmx.h"

static bool read_all(const char* path, std::vector<unsigned char>& out)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize((size_t)n);
    const bool ok = std::fread(out.data(), 1, out.size(), f) == out.size();
    std::fclose(f);
    return ok;
}

static double now_ns()
{
    return std::chrono::duration<double, std::nano>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: abi_threads mesh.bin rays.bin hits_out.bin threads isect\n");
        return 2;
    }
    std::vector<unsigned char> mb, rb;
    if (!read_all(argv[1], mb) || !read_all(argv[2], rb)) {
        std::fprintf(stderr, "cannot read inputs\n");
        return 2;
    }
    uint32_t nvert, ntri, n;
    std::memcpy(&nvert, mb.data(), 4);
    std::memcpy(&ntri, mb.data() + 4, 4);
    std::memcpy(&n, rb.data(), 4);
    std::vector<float> verts(3 * (size_t)nvert);
    std::vector<uint32_t> tris(3 * (size_t)ntri);
    std::memcpy(verts.data(), mb.data() + 8, verts.size() * 4);
    std::memcpy(tris.data(), mb.data() + 8 + verts.size() * 4, tris.size() * 4);
    std::vector<lmx_ray> rays(n);
    std::memcpy(rays.data(), rb.data() + 4, (size_t)n * sizeof(lmx_ray));
    const int nthreads = std::atoi(argv[4]);
    const lmx_build_opts opts = {std::atoi(argv[5]), 0};

    lmx_accel* accel = nullptr;
    if (lmx_accel_build(verts.data(), nvert, tris.data(), ntri, &opts, &accel) != LMX_OK) {
        std::fprintf(stderr, "build: %s\n", lmx_last_error());
        return 1;
    }

    // one batched call on this thread: the reference result
    std::vector<lmx_hit> ref(n), one(n), thr(n);
    double t0 = now_ns();
    if (lmx_accel_intersect_host(accel, rays.data(), n, ref.data(), 1) != LMX_OK) {
        std::fprintf(stderr, "batched: %s\n", lmx_last_error());
        return 1;
    }
    const double ns_batched = (now_ns() - t0) / n;

    // one call per ray on this thread: the per-call cost of the boundary
    t0 = now_ns();
    for (uint32_t i = 0; i < n; i++)
        if (lmx_accel_intersect_host(accel, &rays[i], 1, &one[i], 1) != LMX_OK) {
            std::fprintf(stderr, "per-ray: %s\n", lmx_last_error());
            return 1;
        }
    const double ns_percall = (now_ns() - t0) / n;

    // the scheduler's pattern: workers take grains of 10000 and call the
    // scalar query per ray, all at once on one accel
    const uint32_t grain = 10000;
    std::atomic<uint32_t> next(0);
    std::atomic<int> errors(0);
    auto worker = [&]() {
        while (true) {
            const uint32_t b = next.fetch_add(grain);
            if (b >= n) break;
            const uint32_t e = b + grain < n ? b + grain : n;
            for (uint32_t i = b; i < e; i++)
                if (lmx_accel_intersect_host(accel, &rays[i], 1, &thr[i], 0) != LMX_OK) errors++;
        }
    };
    t0 = now_ns();
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    const double ms_threads = (now_ns() - t0) * 1e-6;

    uint32_t diff_one = 0, diff_thr = 0, hits = 0;
    for (uint32_t i = 0; i < n; i++) {
        diff_one += std::memcmp(&one[i], &ref[i], sizeof(lmx_hit)) != 0;
        diff_thr += std::memcmp(&thr[i], &ref[i], sizeof(lmx_hit)) != 0;
        hits += ref[i].id != LMX_MISS;
    }
    FILE* f = std::fopen(argv[3], "wb");
    if (!f || std::fwrite(thr.data(), sizeof(lmx_hit), n, f) != n) {
        std::fprintf(stderr, "cannot write %s\n", argv[3]);
        return 2;
    }
    std::fclose(f);
    lmx_accel_free(accel);
    std::printf("{\"rays\": %u, \"hits\": %u, \"threads\": %d, \"ns_per_ray_batched\": %.1f, "
                "\"ns_per_ray_percall\": %.1f, \"ms_threaded\": %.3f, \"diff_percall\": %u, "
                "\"diff_threaded\": %u, \"errors\": %d}\n",
                n, hits, nthreads, ns_batched, ns_percall, ms_threads, diff_one, diff_thr, errors.load());
    return (diff_one == 0 && diff_thr == 0 && errors.load() == 0) ? 0 : 1;
}
