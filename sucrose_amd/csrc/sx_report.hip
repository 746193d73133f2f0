// Reporters (reporters.f90): state-data lines, LAMMPS-style trajectory
// frames, configuration dumps and checkpoints, written from the device state.
//
// Text formats.  Every line is built with the reference's edit descriptors as
// its Fortran runtime prints them:
//   I0         minimal-width integer;
//   F0.d       minimal width, d decimals, correctly rounded (ties to even on
//              the exact binary value), no zero before the decimal point when
//              |x| < 1 (".500", "-.062"), the sign kept for a negative value
//              that rounds to zero ("-.000"), NaN / Inf / -Inf spelled so;
//   nX         spaces between items only: an X at the end of a format writes
//              nothing;
//   list-directed integer output (write(unit, *) k): one blank, then I0;
//   an empty list-directed write: an empty line.
//
// Device feed.  sx_report copies the fields a frame needs into a device
// snapshot slot on the context's stream (stream-ordered after the work queued
// before it) and returns; a host thread per reporter waits for the snapshot,
// copies it to pinned host memory on a stream of its own, formats and writes
// it.  Two slots per reporter: a third report waits until the oldest frame
// is on its way to the file.
#include <cerrno>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sx_internal.h"

namespace {

// ---------------------------------------------------------------------------
// edit descriptors
// ---------------------------------------------------------------------------
void put_f0(std::string &o, double x, int d)
{
    if (std::isnan(x)) { o += "NaN"; return; }
    if (std::isinf(x)) { o += x < 0 ? "-Inf" : "Inf"; return; }
    char b[400];
    int len = std::snprintf(b, sizeof b, "%.*f", d, x);
    if (b[0] == '-' && b[1] == '0' && b[2] == '.') {
        o += '-';
        o.append(b + 2, len - 2);
    } else if (b[0] == '0' && b[1] == '.') {
        o.append(b + 1, len - 1);
    } else {
        o.append(b, len);
    }
}

void put_i0(std::string &o, long v) { o += std::to_string(v); }

// items joined by two blanks (the "2x" between items of a format)
struct Line {
    std::string &o;
    bool first = true;
    explicit Line(std::string &o_) : o(o_) {}
    void sep() { if (!first) o += "  "; first = false; }
    void f(double x, int d) { sep(); put_f0(o, x, d); }
    void i(long v) { sep(); put_i0(o, v); }
    void end() { o += '\n'; }
};

// the fields each frame reads (besides scalars, box, types and bonds)
const int TRAJ_FIELDS[] = {SX_SIG_X, SX_SIG_Y, SX_SIG_Z, SX_RX, SX_RY, SX_RZ, SX_FX, SX_FY, SX_FZ,
                           SX_QW, SX_QX, SX_QY, SX_QZ};
const int CONF_FIELDS[] = {SX_RX, SX_RY, SX_RZ, SX_VX, SX_VY, SX_VZ, SX_LX, SX_LY, SX_LZ,
                           SX_QW, SX_QX, SX_QY, SX_QZ};
const int CHK_FIELDS[] = {SX_MASS, SX_CHARGE, SX_SIG_X, SX_SIG_Y, SX_SIG_Z, SX_EPS_X, SX_EPS_Y, SX_EPS_Z,
                          SX_IXX, SX_IYY, SX_IZZ, SX_RX, SX_RY, SX_RZ, SX_VX, SX_VY, SX_VZ,
                          SX_LX, SX_LY, SX_LZ, SX_QW, SX_QX, SX_QY, SX_QZ};

int kind_fields(int kind, const int **list)
{
    switch (kind) {
    case SX_REPORT_TRAJECTORY: *list = TRAJ_FIELDS; return (int)(sizeof TRAJ_FIELDS / sizeof(int));
    case SX_REPORT_CONFIG: *list = CONF_FIELDS; return (int)(sizeof CONF_FIELDS / sizeof(int));
    case SX_REPORT_CHECKPOINT: *list = CHK_FIELDS; return (int)(sizeof CHK_FIELDS / sizeof(int));
    default: *list = nullptr; return 0;
    }
}

struct FrameIn {
    int n = 0, step = 0, nc = 0, config_index = 0;
    const double *const *f = nullptr;  // SX_NFIELDS pointers (only the frame's fields are read)
    const int *rtype = nullptr;
    const double *scal = nullptr;      // SX_NSCALARS
    const double *box = nullptr;
    const int *bbi = nullptr, *bbj = nullptr;
    const double *bblen = nullptr;
};

const char *STATE_HEADER =
    "Step  Potential_Energy  Kinetic_Energy  Total_Energy  Temperature  Pressure  Density\n";

// report_state_data (reporters.f90:20-36)
void fmt_state(std::string &o, const FrameIn &F)
{
    double N = (double)F.n;
    double pot_av = F.scal[SX_POT_ENG] / N, kin_av = F.scal[SX_KIN_ENG] / N, rot_av = F.scal[SX_KIN_ROT] / N;
    double kin_tot = kin_av + rot_av;
    double tot_eng = pot_av + kin_tot;
    Line L(o);
    L.i(F.step);
    L.f(pot_av, 10); L.f(kin_tot, 10); L.f(tot_eng, 10);
    L.f(F.scal[SX_TEMP], 10); L.f(F.scal[SX_PRESSURE], 10); L.f(F.scal[SX_DENSITY], 10);
    L.end();
}

// report_trajectory (reporters.f90:65-91); the vx vy vz columns carry the
// body z axis (UX, UY, UZ) as in the reference
void fmt_traj(std::string &o, const FrameIn &F)
{
    o += "ITEM: TIMESTEP\n ";
    put_i0(o, F.step);
    o += "\nITEM: NUMBER OF ATOMS\n ";
    put_i0(o, F.n);
    o += "\nITEM: BOX BOUNDS pp pp pp\n";
    for (int d = 0; d < 3; ++d) {
        o += ' ';
        Line L(o);
        L.f(-F.box[d] / 2.0, 3);
        L.f(F.box[d] / 2.0, 3);
        L.end();
    }
    o += "ITEM: ATOMS id type shapex shapey shapez x y z vx vy vz fx fy fz quatw quati quatj quatk\n";
    const double *const *f = F.f;
    for (int i = 0; i < F.n; ++i) {
        double qw = f[SX_QW][i], qx = f[SX_QX][i], qy = f[SX_QY][i], qz = f[SX_QZ][i];
        double ux = 2.0 * (qx * qz + qw * qy);
        double uy = 2.0 * (qy * qz - qw * qx);
        double uz = qw * qw - qx * qx - qy * qy + qz * qz;
        o += ' ';
        Line L(o);
        L.i(i + 1);
        L.i(F.rtype ? F.rtype[i] : -1);
        L.f(f[SX_SIG_X][i], 3); L.f(f[SX_SIG_Y][i], 3); L.f(f[SX_SIG_Z][i], 3);
        L.f(f[SX_RX][i], 3); L.f(f[SX_RY][i], 3); L.f(f[SX_RZ][i], 3);
        L.f(ux, 3); L.f(uy, 3); L.f(uz, 3);
        L.f(f[SX_FX][i], 3); L.f(f[SX_FY][i], 3); L.f(f[SX_FZ][i], 3);
        L.f(qw, 3); L.f(qx, 3); L.f(qy, 3); L.f(qz, 3);
        L.end();
    }
}

// report_confwriter (reporters.f90:120-129)
void fmt_conf(std::string &o, const FrameIn &F)
{
    const double *const *f = F.f;
    for (int i = 0; i < F.n; ++i) {
        Line L(o);
        for (int k : CONF_FIELDS) L.f(f[k][i], 3);
        L.end();
    }
}

// report_checkpoint (reporters.f90:160-186)
void fmt_chk(std::string &o, const FrameIn &F)
{
    o += "ATOMS ";
    put_i0(o, F.n);
    o += "\n\nBONDS ";
    put_i0(o, F.nc);
    o += "\n\nBOX ";
    {
        Line L(o);
        L.f(F.box[0], 10); L.f(F.box[1], 10); L.f(F.box[2], 10);
        L.end();
    }
    o += "\nATOM: ID TYPE MASS CHARGE SHAPEX SHAPEY SHAPEZ ESHAPEX ESHAPEY ESHAPEZ IXX IYY IZZ RX RY RZ VX VY VZ "
         "LX LY LZ QW QX QY QZ\n\n";
    const double *const *f = F.f;
    for (int i = 0; i < F.n; ++i) {
        o += ' ';
        Line L(o);
        L.i(i + 1);
        L.i(F.rtype ? F.rtype[i] : -1);
        for (int k : CHK_FIELDS) L.f(f[k][i], 10);
        L.end();
    }
    o += "\nBOND: ID I J LEN\n\n";
    for (int b = 0; b < F.nc; ++b) {
        o += ' ';
        Line L(o);
        L.i(b + 1);
        L.i(F.bbi[b]);
        L.i(F.bbj[b]);
        L.f(F.bblen[b], 10);
        L.end();
    }
    o += "\n ";
    put_i0(o, F.config_index);
    o += '\n';
}

void format_frame(int kind, std::string &o, const FrameIn &F)
{
    switch (kind) {
    case SX_REPORT_STATE_DATA: fmt_state(o, F); break;
    case SX_REPORT_TRAJECTORY: fmt_traj(o, F); break;
    case SX_REPORT_CONFIG: fmt_conf(o, F); break;
    case SX_REPORT_CHECKPOINT: fmt_chk(o, F); break;
    }
}

// ---------------------------------------------------------------------------
// list-directed input (load_checkpoint)
// ---------------------------------------------------------------------------
struct ListReader {
    FILE *f;
    std::vector<std::string> tok;   // the rest of the current record
    size_t pos = 0;
    bool eof = false;
    explicit ListReader(FILE *f_) : f(f_) {}
    bool next_record()
    {
        tok.clear();
        pos = 0;
        std::string line;
        int ch;
        bool any = false;
        while ((ch = std::fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            line += (char)ch;
        }
        if (!any) { eof = true; return false; }
        size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == ',' || line[i] == '\r')) ++i;
            size_t j = i;
            while (j < line.size() && !(line[j] == ' ' || line[j] == '\t' || line[j] == ',' || line[j] == '\r')) ++j;
            if (j > i) tok.push_back(line.substr(i, j - i));
            i = j;
        }
        return true;
    }
    // one read statement: starts at a new record, continues over records
    // until every item is read
    std::vector<std::string> read(int nitems)
    {
        std::vector<std::string> out;
        if (!next_record()) throw sx_error{SX_ERR_ARG, "checkpoint file ends early"};
        while ((int)out.size() < nitems) {
            if (pos < tok.size()) {
                out.push_back(tok[pos++]);
            } else if (!next_record()) {
                throw sx_error{SX_ERR_ARG, "checkpoint file ends early"};
            }
        }
        return out;
    }
};

double to_real(const std::string &t)
{
    std::string s = t;
    for (char &ch : s)
        if (ch == 'd' || ch == 'D') ch = 'E';
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0') throw sx_error{SX_ERR_ARG, "bad real in checkpoint: " + t};
    return v;
}

long to_int(const std::string &t)
{
    char *end = nullptr;
    long v = std::strtol(t.c_str(), &end, 10);
    if (end == t.c_str() || *end != '\0') throw sx_error{SX_ERR_ARG, "bad integer in checkpoint: " + t};
    return v;
}

struct Checkpoint {
    int n = 0, nc = 0, config_index = 0;
    double box[3] = {0, 0, 0};
    std::vector<int> rtype;
    std::vector<double> f[SX_NFIELDS];   // only the checkpoint fields are filled
    std::vector<int> bbi, bbj;
    std::vector<double> bblen;
};

// load_checkpoint (reporters.f90:188-221): the mass of each line goes to the
// atom the line names (set_atom_mass(J, ...)); every other column to the
// line's position I; inv_mass and the inverse inertia keep set_num_atoms'
// -1; bond lines go to the bond they name (add_constraint(J, ...))
Checkpoint read_checkpoint(const char *path)
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) throw sx_error{SX_ERR_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno)};
    Checkpoint C;
    try {
        ListReader R(fp);
        C.n = (int)to_int(R.read(2)[1]);
        if (C.n <= 0) throw sx_error{SX_ERR_ARG, "checkpoint has no atoms"};
        R.read(0);
        C.nc = (int)to_int(R.read(2)[1]);
        if (C.nc < 0) throw sx_error{SX_ERR_ARG, "bad bond count"};
        R.read(0);
        auto b = R.read(4);
        for (int d = 0; d < 3; ++d) C.box[d] = to_real(b[1 + d]);
        R.read(0);
        R.read(1);
        R.read(0);
        const double init[SX_NFIELDS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                         -1, -1, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0};
        for (int k = 0; k < SX_NFIELDS; ++k) C.f[k].assign(C.n, init[k]);
        C.rtype.assign(C.n, -1);
        for (int i = 0; i < C.n; ++i) {
            auto t = R.read(26);
            long j = to_int(t[0]);
            if (j < 1 || j > C.n) throw sx_error{SX_ERR_ARG, "Error: Atom index does not exist in system"};
            C.rtype[i] = (int)to_int(t[1]);
            C.f[SX_MASS][j - 1] = to_real(t[2]);
            for (int k = 1; k < 24; ++k) C.f[CHK_FIELDS[k]][i] = to_real(t[2 + k]);
        }
        R.read(0);
        R.read(1);
        R.read(0);
        C.bbi.assign(C.nc, 0);
        C.bbj.assign(C.nc, 0);
        C.bblen.assign(C.nc, 0.0);
        for (int q = 0; q < C.nc; ++q) {
            auto t = R.read(4);
            long id = to_int(t[0]);
            if (id < 1 || id > C.nc) throw sx_error{SX_ERR_ARG, "Error: Constraint index does not exist in system"};
            C.bbi[id - 1] = (int)to_int(t[1]);
            C.bbj[id - 1] = (int)to_int(t[2]);
            C.bblen[id - 1] = to_real(t[3]);
        }
        R.read(0);
        C.config_index = (int)to_int(R.read(1)[0]);
    } catch (...) {
        std::fclose(fp);
        throw;
    }
    std::fclose(fp);
    return C;
}

// the state a checkpoint does not hold (thermostat chain, target
// temperature, Langevin step counter, periodic flag), in a sidecar file
std::string sidecar_path(const std::string &p) { return p + ".state"; }

}  // namespace

// ---------------------------------------------------------------------------
// the reporter object
// ---------------------------------------------------------------------------
struct sx_reporter {
    sx_ctx *c = nullptr;
    int kind = 0;
    std::string path;
    FILE *fp = nullptr;
    int config_index = 0;
    std::string err;         // the last failed call's message
    std::string async_err;   // the first failure of the writer thread, reported by flush and close
    // device feed
    int n = 0, nf = 0;
    const int *fields = nullptr;
    double *d_snap[2] = {nullptr, nullptr};    // nf * n field values, then the scalars and the step
    double *h_snap[2] = {nullptr, nullptr};    // pinned
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
    bool busy[2] = {false, false};
    struct Frame {
        int slot, step, config_index;
        double box[3];
        int periodic;
        std::vector<int> rtype, bbi, bbj;
        std::vector<double> bblen;
    };
    std::deque<Frame> queue;
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    bool stop = false, working = false;

    long snap_len() const { return (long)nf * n + SX_NSCALARS + 1; }

    void write(const std::string &s)
    {
        if (s.empty()) return;
        if (std::fwrite(s.data(), 1, s.size(), fp) != s.size()) throw sx_error{SX_ERR_ARG, "write failed: " + path};
    }

    void write_frame_host(const FrameIn &F)
    {
        std::string o;
        o.reserve(kind == SX_REPORT_STATE_DATA ? 128 : (size_t)F.n * (kind == SX_REPORT_CHECKPOINT ? 330 : 110) + 512);
        format_frame(kind, o, F);
        write(o);
        std::fflush(fp);
    }

    void run()
    {
        hipSetDevice(c->device);
        for (;;) {
            Frame fr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !queue.empty(); });
                if (queue.empty()) return;
                fr = std::move(queue.front());
                queue.pop_front();
                working = true;
            }
            try {
                int s = fr.slot;
                HIPCHK(hipEventSynchronize(ev[s]));
                HIPCHK(hipMemcpyAsync(h_snap[s], d_snap[s], sizeof(double) * snap_len(), hipMemcpyDeviceToHost,
                                      copy_stream));
                HIPCHK(hipStreamSynchronize(copy_stream));
                const double *ptr[SX_NFIELDS] = {nullptr};
                for (int k = 0; k < nf; ++k) ptr[fields[k]] = h_snap[s] + (long)k * n;
                const double *scal = h_snap[s] + (long)nf * n;
                FrameIn F;
                F.n = n; F.step = fr.step; F.nc = (int)fr.bbi.size(); F.config_index = fr.config_index;
                F.f = ptr; F.rtype = fr.rtype.data(); F.scal = scal; F.box = fr.box;
                F.bbi = fr.bbi.data(); F.bbj = fr.bbj.data(); F.bblen = fr.bblen.data();
                write_frame_host(F);
                if (kind == SX_REPORT_CHECKPOINT) write_sidecar(scal, (long)scal[SX_NSCALARS], fr.periodic);
            } catch (const sx_error &e) {
                std::lock_guard<std::mutex> lk(mu);
                if (async_err.empty()) async_err = e.msg;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                busy[fr.slot] = false;
                working = false;
            }
            cv.notify_all();
        }
    }

    void write_sidecar(const double *scal, long lgv_step, int periodic)
    {
        std::string p = sidecar_path(path);
        FILE *f = std::fopen(p.c_str(), "w");
        if (!f) throw sx_error{SX_ERR_ARG, "cannot write " + p};
        std::fprintf(f, "TARGET_TEMP %.17g\nDOF %.17g\n", scal[SX_TARGET_TEMP], scal[SX_DOF]);
        std::fprintf(f, "HB %.17g %.17g %.17g\n", scal[SX_HB0], scal[SX_HB1], scal[SX_HB2]);
        std::fprintf(f, "ETA %.17g %.17g %.17g\n", scal[SX_ETA0], scal[SX_ETA1], scal[SX_ETA2]);
        std::fprintf(f, "P_ETA %.17g %.17g %.17g\n", scal[SX_PETA0], scal[SX_PETA1], scal[SX_PETA2]);
        std::fprintf(f, "LANGEVIN_STEP %ld\nPERIODIC %d\n", lgv_step, periodic);
        std::fclose(f);
    }

    void drain()
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return queue.empty() && !working; });
    }
};

namespace {

int rguard(sx_reporter *r, const std::function<void()> &fn)
{
    if (!r) return SX_ERR_ARG;
    try {
        if (r->c) HIPCHK(hipSetDevice(r->c->device));
        fn();
        return SX_OK;
    } catch (const sx_error &e) {
        std::lock_guard<std::mutex> lk(r->mu);
        r->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(r->mu);
        r->err = e.what();
        return SX_ERR_HIP;
    }
}

void free_feed(sx_reporter *r)
{
    for (int s = 0; s < 2; ++s) {
        if (r->d_snap[s]) hipFree(r->d_snap[s]);
        if (r->h_snap[s]) hipHostFree(r->h_snap[s]);
        if (r->ev[s]) hipEventDestroy(r->ev[s]);
        r->d_snap[s] = r->h_snap[s] = nullptr;
        r->ev[s] = nullptr;
    }
}

__global__ void k_copy_step(const int *dstep, double *out) { *out = dstep ? (double)*dstep : 0.0; }

}  // namespace

extern "C" {

int sx_reporter_open(sx_ctx *c, int kind, const char *filename, sx_reporter **out)
{
    if (!out || !filename) return SX_ERR_ARG;
    *out = nullptr;
    if (kind < SX_REPORT_STATE_DATA || kind > SX_REPORT_CHECKPOINT) return SX_ERR_ARG;
    sx_reporter *r = new sx_reporter();
    r->c = c;
    r->kind = kind;
    r->path = filename;
    int rc = rguard(r, [&] {
        r->fp = std::fopen(filename, "w");
        if (!r->fp) throw sx_error{SX_ERR_ARG, std::string("cannot open ") + filename + ": " + std::strerror(errno)};
        if (kind == SX_REPORT_STATE_DATA) {   // open_state_data_file writes the header (reporters.f90:16)
            r->write(STATE_HEADER);
            std::fflush(r->fp);
        }
        if (c) {
            r->nf = kind_fields(kind, &r->fields);
            HIPCHK(hipStreamCreateWithFlags(&r->copy_stream, hipStreamNonBlocking));
            r->worker = std::thread([r] { r->run(); });
        }
    });
    if (rc != SX_OK) {
        if (c) c->err = r->err;
        if (r->fp) std::fclose(r->fp);
        if (r->copy_stream) hipStreamDestroy(r->copy_stream);
        delete r;
        return rc;
    }
    *out = r;
    return SX_OK;
}

int sx_report(sx_reporter *r, int step)
{
    return rguard(r, [&] {
        sx_ctx *c = r->c;
        if (!c) throw sx_error{SX_ERR_STATE, "reporter has no context (use sx_report_host)"};
        need_system(c);
        if (c->n != r->n) {   // first frame or a new system size: (re)allocate the slots
            r->drain();
            free_feed(r);
            r->n = c->n;
            for (int s = 0; s < 2; ++s) {
                HIPCHK(hipMalloc(&r->d_snap[s], sizeof(double) * r->snap_len()));
                HIPCHK(hipHostMalloc(&r->h_snap[s], sizeof(double) * r->snap_len(), hipHostMallocDefault));
                HIPCHK(hipEventCreateWithFlags(&r->ev[s], hipEventDisableTiming));
            }
        }
        int s;
        {
            std::unique_lock<std::mutex> lk(r->mu);
            r->cv.wait(lk, [&] { return !r->busy[0] || !r->busy[1]; });
            s = r->busy[0] ? 1 : 0;
            r->busy[s] = true;
        }
        try {
            long n = r->n;
            for (int k = 0; k < r->nf; ++k)
                HIPCHK(hipMemcpyAsync(r->d_snap[s] + k * n, c->d_fields + (long)r->fields[k] * c->ld,
                                      sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(r->d_snap[s] + r->nf * n, c->d_scal, sizeof(double) * SX_NSCALARS,
                                  hipMemcpyDeviceToDevice, c->stream));
            hipLaunchKernelGGL(k_copy_step, dim3(1), dim3(1), 0, c->stream, (const int *)c->d_step,
                               r->d_snap[s] + r->nf * n + SX_NSCALARS);
            HIPCHK(hipEventRecord(r->ev[s], c->stream));
        } catch (...) {
            std::lock_guard<std::mutex> lk(r->mu);
            r->busy[s] = false;
            throw;
        }
        sx_reporter::Frame fr;
        fr.slot = s;
        fr.step = step;
        fr.config_index = r->config_index;
        for (int d = 0; d < 3; ++d) fr.box[d] = c->box[d];
        fr.periodic = c->periodic;
        fr.rtype = c->h_rtype;
        if (r->kind == SX_REPORT_CHECKPOINT) {
            fr.bbi = c->h_bbi;
            fr.bbj = c->h_bbj;
            fr.bblen = c->h_bblen;
        }
        {
            std::lock_guard<std::mutex> lk(r->mu);
            r->queue.push_back(std::move(fr));
        }
        r->cv.notify_all();
    });
}

int sx_report_host(sx_reporter *r, int step, int n, const double *const *fields, const int *rtype,
                   const double *scal, const double box[3], int nc, const int *bbi1, const int *bbj1,
                   const double *bblen)
{
    return rguard(r, [&] {
        if (n <= 0 || !fields || !scal || !box || nc < 0 || (nc > 0 && (!bbi1 || !bbj1 || !bblen)))
            throw sx_error{SX_ERR_ARG, "bad frame"};
        const int *list = nullptr;
        int nf = kind_fields(r->kind, &list);
        for (int k = 0; k < nf; ++k)
            if (!fields[list[k]]) throw sx_error{SX_ERR_ARG, "a field the frame needs is missing"};
        if (r->c) r->drain();
        FrameIn F;
        F.n = n; F.step = step; F.nc = nc; F.config_index = r->config_index;
        F.f = fields; F.rtype = rtype; F.scal = scal; F.box = box;
        F.bbi = bbi1; F.bbj = bbj1; F.bblen = bblen;
        r->write_frame_host(F);
    });
}

int sx_reporter_set_config_index(sx_reporter *r, int config_index)
{
    return rguard(r, [&] { r->config_index = config_index; });
}

int sx_reporter_flush(sx_reporter *r)
{
    return rguard(r, [&] {
        if (r->c) r->drain();
        std::fflush(r->fp);
        std::string e;
        {
            std::lock_guard<std::mutex> lk(r->mu);
            e = r->async_err;
        }
        if (!e.empty()) throw sx_error{SX_ERR_ARG, e};
    });
}

const char *sx_reporter_error(const sx_reporter *r) { return r ? r->err.c_str() : "null reporter"; }

int sx_reporter_close(sx_reporter *r)
{
    if (!r) return SX_ERR_ARG;
    int rc = sx_reporter_flush(r);
    if (r->c) {
        {
            std::lock_guard<std::mutex> lk(r->mu);
            r->stop = true;
        }
        r->cv.notify_all();
        if (r->worker.joinable()) r->worker.join();
        hipSetDevice(r->c->device);
        free_feed(r);
        if (r->copy_stream) hipStreamDestroy(r->copy_stream);
    }
    if (r->fp) std::fclose(r->fp);
    delete r;
    return rc;
}

int sx_read_checkpoint_host(const char *filename, int *n, int *nc, double box[3], int *config_index, int cap_n,
                            int cap_nc, double *const *fields, int *rtype, int *bbi1, int *bbj1, double *bblen)
{
    try {
        Checkpoint C = read_checkpoint(filename);
        if (n) *n = C.n;
        if (nc) *nc = C.nc;
        if (box) for (int d = 0; d < 3; ++d) box[d] = C.box[d];
        if (config_index) *config_index = C.config_index;
        if (C.n > cap_n || C.nc > cap_nc) return SX_ERR_CAPACITY;
        if (fields)
            for (int k = 0; k < SX_NFIELDS; ++k)
                if (fields[k]) std::copy(C.f[k].begin(), C.f[k].end(), fields[k]);
        if (rtype) std::copy(C.rtype.begin(), C.rtype.end(), rtype);
        if (bbi1) std::copy(C.bbi.begin(), C.bbi.end(), bbi1);
        if (bbj1) std::copy(C.bbj.begin(), C.bbj.end(), bbj1);
        if (bblen) std::copy(C.bblen.begin(), C.bblen.end(), bblen);
        return SX_OK;
    } catch (const sx_error &e) {
        return e.code;
    }
}

int sx_load_checkpoint(sx_ctx *c, const char *filename, int *config_index)
{
    return guard(c, [&] {
        if (!filename) throw sx_error{SX_ERR_ARG, "null file name"};
        Checkpoint C = read_checkpoint(filename);
        int periodic = c->periodic;
        // set_num_atoms, set_num_constraints, set_box (reporters.f90:196-203)
        int rc = sx_set_num_atoms(c, C.n, 1);
        if (rc != SX_OK) throw sx_error{rc, c->err};
        // the sidecar written beside a checkpoint by sx_report, when present
        std::vector<std::pair<std::string, std::vector<double>>> side;
        if (FILE *f = std::fopen(sidecar_path(filename).c_str(), "r")) {
            char key[64];
            while (std::fscanf(f, "%63s", key) == 1) {
                std::vector<double> v;
                int want = (!std::strcmp(key, "HB") || !std::strcmp(key, "ETA") || !std::strcmp(key, "P_ETA")) ? 3 : 1;
                for (int q = 0; q < want; ++q) {
                    double x;
                    if (std::fscanf(f, "%lf", &x) != 1) throw sx_error{SX_ERR_ARG, "bad sidecar " + sidecar_path(filename)};
                    v.push_back(x);
                }
                side.emplace_back(key, v);
            }
            std::fclose(f);
        }
        for (auto &kv : side)
            if (kv.first == "PERIODIC") periodic = (int)kv.second[0];
        rc = sx_set_box(c, C.box, periodic);
        if (rc != SX_OK) throw sx_error{rc, c->err};
        for (int k = 0; k < SX_NFIELDS; ++k)
            HIPCHK(hipMemcpyAsync(c->d_fields + (long)k * c->ld, C.f[k].data(), sizeof(double) * C.n,
                                  hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->h_rtype = C.rtype;
        if (C.nc > 0) {
            rc = sx_set_constraints(c, C.nc, C.bbi.data(), C.bbj.data(), C.bblen.data());
            if (rc != SX_OK) throw sx_error{rc, c->err};
        }
        sx_sync_scalars_to_host(c);
        for (auto &kv : side) {
            const auto &v = kv.second;
            if (kv.first == "TARGET_TEMP") { c->h_scal[SX_TARGET_TEMP] = v[0]; c->have_target_temp = true; }
            else if (kv.first == "DOF") c->h_scal[SX_DOF] = v[0];
            else if (kv.first == "HB") for (int q = 0; q < 3; ++q) c->h_scal[SX_HB0 + q] = v[q];
            else if (kv.first == "ETA") for (int q = 0; q < 3; ++q) c->h_scal[SX_ETA0 + q] = v[q];
            else if (kv.first == "P_ETA") for (int q = 0; q < 3; ++q) c->h_scal[SX_PETA0 + q] = v[q];
            else if (kv.first == "LANGEVIN_STEP") {
                if (!c->d_step) HIPCHK(hipMalloc(&c->d_step, sizeof(int)));
                int st = (int)v[0];
                HIPCHK(hipMemcpy(c->d_step, &st, sizeof(int), hipMemcpyHostToDevice));
            }
        }
        sx_sync_scalars_to_device(c);
        c->have_graph = false;
        if (config_index) *config_index = C.config_index;
    });
}

}  // extern "C"
