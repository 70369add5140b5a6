// Drives libfvm2d_gpu's host code through bench.py's call sequences (main
// leg, 8-angle batches, 48 requests one after another and batched, a serving
// engine's join/leave pattern, error paths and stale handles) against the HIP
// stand-in (hip_stand_in.cpp), under AddressSanitizer.
//
//   drive_host DIR [REPS]     the sequences over the meshes in DIR
//   drive_host --plant-scope  a planted mistake: an asynchronous copy from a
//                             host buffer whose scope ends before the stream
//                             is synchronised (must be reported)
#include "fvm2d_gpu.h"
#include <hip/hip_runtime_api.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

extern "C" void hipstub_report(void);

struct Req {
    std::string name;
    int np = 0, nt = 0;
    fvm_params par{};
    std::vector<double> coord;
    std::vector<int> elem, ptype;
    // geometric() outputs
    int dims[12]{};
    std::vector<int> edge, edneigh, iedge, beindx, spts;
    std::vector<double> ds, dsb, elarea, cvarea, mcarea, drmin, wd, wd1;
    fvm_mesh m{};
};

static void check(int rc, const char *what, bool allow_err = false)
{
    if (rc < 0 && !allow_err) {
        fprintf(stderr, "%s -> %d (%s)\n", what, rc, fvm_status_string(rc));
        exit(2);
    }
}

static Req *load(const std::string &path)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) { perror(path.c_str()); exit(2); }
    Req *r = new Req;
    r->name = path;
    int hd[2];
    if (fread(hd, 4, 2, f) != 2) exit(3);
    r->np = hd[0];
    r->nt = hd[1];
    if (fread(&r->par, sizeof(fvm_params), 1, f) != 1) exit(3);
    r->coord.resize(2 * r->np);
    r->elem.resize(3 * r->nt);
    r->ptype.resize(r->np);
    if (fread(r->coord.data(), 8, r->coord.size(), f) != r->coord.size()) exit(3);
    if (fread(r->elem.data(), 4, r->elem.size(), f) != r->elem.size()) exit(3);
    if (fread(r->ptype.data(), 4, r->ptype.size(), f) != r->ptype.size()) exit(3);
    fclose(f);
    return r;
}

static void geometric(Req *r)
{
    const int me = 3 * r->nt + 3, np = r->np;
    r->edge.assign(2 * me, 0);
    r->edneigh.assign(2 * me, 0);
    r->iedge.assign(me, 0);
    r->ds.assign(2 * me, 0.0);
    r->dsb.assign(2 * np, 0.0);
    r->elarea.assign(r->nt, 0.0);
    r->cvarea.assign(np, 0.0);
    r->mcarea.assign(np, 0.0);
    r->drmin.assign(np, 0.0);
    r->wd.assign(np, 0.0);
    r->wd1.assign(np, 0.0);
    r->beindx.assign(me, 0);
    r->spts.assign(np, 0);
    check(fvm_geometric(np, r->nt, r->coord.data(), r->elem.data(), r->ptype.data(), r->par.cell_type,
                        1, r->par.sa_compat, me, r->edge.data(), r->edneigh.data(), r->iedge.data(),
                        r->ds.data(), r->dsb.data(), r->elarea.data(), r->cvarea.data(),
                        r->mcarea.data(), r->drmin.data(), r->wd.data(), r->wd1.data(),
                        r->beindx.data(), r->spts.data(), r->dims),
          "fvm_geometric");
    const int *d = r->dims;
    fvm_mesh &m = r->m;
    m.np = d[0]; m.nt = d[1]; m.ne = d[2]; m.netot = d[3]; m.nsp = d[4];
    m.nswe1 = d[7]; m.nswe2 = d[8]; m.nffe1 = d[9]; m.nffe2 = d[10];
    // trim to the sizes geometric() reports, as the Python host does, so
    // that reads past them are caught
    r->edge.resize(2 * m.netot);
    r->iedge.resize(m.ne);
    r->ds.resize(2 * m.netot);
    r->beindx.resize(d[6]);
    r->spts.resize(m.nsp);
    r->edge.shrink_to_fit(); r->iedge.shrink_to_fit(); r->ds.shrink_to_fit();
    r->beindx.shrink_to_fit(); r->spts.shrink_to_fit();
    m.coord = r->coord.data(); m.elem = r->elem.data(); m.edge = r->edge.data();
    m.iedge = r->iedge.data(); m.beindx = r->beindx.data(); m.ptype = r->ptype.data();
    m.spts = r->spts.data(); m.ds = r->ds.data(); m.elarea = r->elarea.data();
    m.cvarea = r->cvarea.data(); m.mcarea = r->mcarea.data(); m.drmin = r->drmin.data();
    m.wd = r->wd.data();
}

static void walldist(Req *r)
{
    std::vector<double> wd(r->np), wd1(std::max(r->m.nsp, 1));
    check(fvm_gpu_wall_dist(r->np, r->coord.data(), r->ptype.data(), r->m.netot, r->edge.data(), 0,
                            wd.data(), wd1.data()),
          "wall_dist");
    r->wd = wd;
    r->m.wd = r->wd.data();
}

static fvm_ctx *create(Req *r, fvm_params p)
{
    fvm_ctx *x = nullptr;
    check(fvm_gpu_create(&p, &r->m, &x), "create");
    return x;
}

static void iterate(fvm_ctx *x, int n, bool hist = true)
{
    std::vector<double> h(6 * (size_t)std::max(n, 1));
    int nd = 0, ei = 0;
    check(fvm_gpu_iterate(x, n, hist ? h.data() : nullptr, &nd, &ei), "iterate", true);
}

static int plant_scope()
{
    hipStream_t s = nullptr;
    void *d = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipMalloc(&d, 64) != hipSuccess) return 2;
    {
        double block_scoped[8] = {1, 2, 3, 4, 5, 6, 7, 8};
        (void)hipMemcpyAsync(d, block_scoped, sizeof block_scoped, hipMemcpyHostToDevice, s);
    }
    (void)hipStreamSynchronize(s);      // the copy runs here, after the buffer's scope ended
    (void)hipFree(d);
    (void)hipStreamDestroy(s);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && !strcmp(argv[1], "--plant-scope")) return plant_scope();
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const int reps = argc > 2 ? atoi(argv[2]) : 1;
    Req *mainq = load(dir + "/main_513x129.bin");
    geometric(mainq);
    std::vector<Req *> reqs;
    for (int i = 0; i < 48; ++i) {
        char b[64];
        snprintf(b, sizeof b, "/req_%02d.bin", i);
        reqs.push_back(load(dir + b));
        geometric(reqs.back());
    }
    fprintf(stderr, "loaded; main np %d\n", mainq->np);

    for (int rep = 0; rep < reps; ++rep) {
        // main leg
        {
            fvm_ctx *s = create(mainq, mainq->par);
            iterate(s, 50);
            iterate(s, 200);
            double ms[16];
            const char *names[16];
            fvm_gpu_time_kernels(s, 20, ms, names, 16);
            double t = 0;
            check(fvm_gpu_time_stage(s, 100, &t), "time_stage", true);
            std::vector<double> prim(4 * mainq->np), nut(mainq->np), qx(4 * mainq->np), qy(4 * mainq->np),
                mu(mainq->np), res(4 * mainq->np);
            check(fvm_gpu_get_state(s, prim.data(), nut.data(), qx.data(), qy.data(), mu.data()), "get_state");
            check(fvm_gpu_set_state(s, prim.data(), nullptr), "set_state");
            check(fvm_gpu_residu(s, prim.data(), nullptr, res.data()), "residu", true);
            double mon[6];
            fvm_gpu_monitor(s, mon);
            long long info[8];
            fvm_gpu_layout_info(s, info);
            (void)fvm_gpu_footprint(s);
            check(fvm_gpu_tune(s, 1, 64, 1), "tune", true);
            iterate(s, 10);
            check(fvm_gpu_tune(s, 3, 64, 1), "tune", true);
            iterate(s, 10);
            check(fvm_gpu_destroy(s), "destroy");
            double gbs;
            check(fvm_gpu_stream_copy(1 << 20, 3, &gbs), "stream_copy");
        }
        fprintf(stderr, "main leg done\n");
        // 8 alphas, FP64 and split, one batch
        for (int prec = 0; prec < 2; ++prec) {
            std::vector<fvm_ctx *> sol;
            for (int k = 0; k < 8; ++k) {
                fvm_params p = mainq->par;
                p.aoa_deg = 2.0 * k;
                p.precision = prec;
                sol.push_back(create(mainq, p));
            }
            fvm_batch *b = nullptr;
            check(fvm_batch_create(&b), "batch_create");
            for (auto *x : sol) check(fvm_batch_add(b, x), "batch_add");
            std::vector<int> st(8);
            check(fvm_batch_iterate(b, 50, st.data()), "batch_iterate");
            check(fvm_batch_iterate(b, 200, st.data()), "batch_iterate");
            check(fvm_batch_destroy(b), "batch_destroy");
            for (auto *x : sol) check(fvm_gpu_destroy(x), "destroy");
        }
        fprintf(stderr, "alpha batches done\n");
        // 48 requests, plain FP64 one after another
        for (Req *r : reqs) {
            if (r->par.iflow == 3) walldist(r);
            fvm_params p = r->par;
            p.precision = 0;
            p.patch_size = 0;
            fvm_ctx *x = create(r, p);
            iterate(x, 8);
            double mon[6];
            fvm_gpu_monitor(x, mon);
            iterate(x, 200);
            fvm_gpu_monitor(x, mon);
            check(fvm_gpu_destroy(x), "destroy");
        }
        fprintf(stderr, "plain requests done\n");
        // the same batched, one batch per flow kind, chunks launched together
        {
            std::vector<fvm_ctx *> eu, ra;
            for (Req *r : reqs) (r->par.iflow == 3 ? ra : eu).push_back(create(r, r->par));
            fvm_batch *be = nullptr, *br = nullptr;
            check(fvm_batch_create(&be), "batch_create");
            check(fvm_batch_create(&br), "batch_create");
            for (auto *x : eu) check(fvm_batch_add(be, x), "add");
            for (auto *x : ra) check(fvm_batch_add(br, x), "add");
            std::vector<int> st(64);
            check(fvm_batch_iterate(be, 8, st.data()), "it");
            check(fvm_batch_iterate(br, 8, st.data()), "it");
            for (int k = 0; k < 4; ++k) {
                check(fvm_batch_launch(be, 50), "launch");
                check(fvm_batch_launch(br, 50), "launch");
                check(fvm_batch_wait(be, st.data()), "wait");
                check(fvm_batch_wait(br, st.data()), "wait");
            }
            double mon[6];
            for (auto *x : eu) fvm_gpu_monitor(x, mon);
            check(fvm_batch_destroy(be), "bd");
            check(fvm_batch_destroy(br), "bd");
            for (auto *x : eu) check(fvm_gpu_destroy(x), "destroy");
            for (auto *x : ra) check(fvm_gpu_destroy(x), "destroy");
        }
        fprintf(stderr, "mixed batches done\n");
        // serving engine pattern: members join and leave between chunks,
        // host work while a chunk runs, contexts destroyed after removal
        {
            fvm_batch *be = nullptr;
            check(fvm_batch_create(&be), "batch_create");
            std::vector<fvm_ctx *> live;
            size_t next = 0;
            std::vector<int> st(64);
            for (int step = 0; step < 40; ++step) {
                while (next < reqs.size() && live.size() < 6) {
                    Req *r = reqs[next++];
                    if (r->par.iflow == 3) continue;
                    fvm_ctx *x = create(r, r->par);
                    check(fvm_batch_add(be, x), "add");
                    live.push_back(x);
                }
                if (live.empty()) break;
                check(fvm_batch_launch(be, 16), "launch");
                int spins = 0;
                while (!fvm_batch_query(be) && spins < 4) ++spins;
                // host work while the chunk runs: a monitor read of a member
                // (waits for the batch) and a request's preprocessing
                double mon[6];
                fvm_gpu_monitor(live[0], mon);
                check(fvm_batch_wait(be, st.data()), "wait");
                // retire the oldest two
                for (int k = 0; k < 2 && !live.empty(); ++k) {
                    fvm_ctx *x = live.front();
                    live.erase(live.begin());
                    if (k == 0) check(fvm_batch_remove(be, x), "remove");
                    std::vector<double> prim(4 * 1024 * 1024);
                    check(fvm_gpu_get_state(x, prim.data(), nullptr, nullptr, nullptr, nullptr), "gs");
                    check(fvm_gpu_destroy(x), "destroy");   // k == 1: destroyed while a member
                }
            }
            for (auto *x : live) check(fvm_gpu_destroy(x), "destroy");
            check(fvm_batch_destroy(be), "bd");
        }
        fprintf(stderr, "serving pattern done\n");
        // error paths: a member's own iterate is refused; destroy with a
        // chunk pending; stale handles refused
        {
            fvm_ctx *x = create(reqs[1], reqs[1]->par), *y = create(reqs[2], reqs[2]->par);
            fvm_batch *b = nullptr;
            check(fvm_batch_create(&b), "bc");
            check(fvm_batch_add(b, x), "add");
            int nd, ei;
            if (fvm_gpu_iterate(x, 5, nullptr, &nd, &ei) != FVM_E_INPUT) { fprintf(stderr, "member iterate allowed\n"); return 4; }
            check(fvm_batch_launch(b, 8), "launch");
            check(fvm_batch_add(b, y), "add while pending");
            check(fvm_batch_launch(b, 8), "launch");
            check(fvm_batch_destroy(b), "destroy pending");
            iterate(x, 4);
            check(fvm_gpu_destroy(x), "destroy");
            check(fvm_gpu_destroy(y), "destroy");
            // destroyed handles are refused without being read
            double mon[6];
            if (fvm_gpu_iterate(x, 1, nullptr, &nd, &ei) != FVM_E_INPUT || fvm_gpu_monitor(y, mon) != FVM_E_INPUT ||
                fvm_gpu_destroy(x) != FVM_E_INPUT || fvm_batch_launch(b, 1) != FVM_E_INPUT ||
                fvm_batch_destroy(b) != FVM_E_INPUT || fvm_batch_size(b) != 0 ||
                fvm_batch_add(b, y) != FVM_E_INPUT) {
                fprintf(stderr, "stale handle accepted\n");
                return 5;
            }
            // a batch handle passed as a context and the reverse
            fvm_batch *b2 = nullptr;
            check(fvm_batch_create(&b2), "bc");
            if (fvm_gpu_iterate(reinterpret_cast<fvm_ctx *>(b2), 1, nullptr, &nd, &ei) != FVM_E_INPUT) {
                fprintf(stderr, "batch accepted as a context\n");
                return 6;
            }
            check(fvm_batch_destroy(b2), "bd");
        }
        // no-allocation wall distance on "device" (host) buffers
        {
            Req *r = reqs[0];
            std::vector<double> wd(r->np);
            check(fvm_gpu_wall_dist_noalloc(r->np, r->coord.data(), r->ptype.data(), r->m.netot,
                                            r->edge.data(), 0, wd.data(), nullptr),
                  "noalloc");
        }
        fprintf(stderr, "rep %d done\n", rep);
    }
    hipstub_report();
    for (Req *r : reqs) delete r;
    delete mainq;
    return 0;
}
