// mdg_io.cpp -- MDBas input and output formats and the host-side set-up
// steps, so a host can run a reference deck (SIMU / CONF / FORF or REST)
// through libmdgpu and write the reference's PROP / REST / RCON / DCD
// files.  Host code only; no device work.
//
//   readers  io.c:29-726 (SIMU), 728-783 (CONF), 785-870 (REST),
//            872-1199 (FORF)
//   writers  io.c:1201-1228 (RCON), 1230-1274 (PROP), 1276-1369 (REST),
//            1371-1460 (DCD header), 1462-1522 (DCD frame)
//   set-up   init.c:102-109 (Coulomb constant), 248-616 (setup),
//            627-833 (init_vel with the caller's Gaussian deviates),
//            840-908 (init_box); utils.c:162-218 (image_update),
//            502-515 (kinetic), 551-582 (degrees of freedom);
//            spme.c:32-73 (Ewald width and mesh size)
//
// Every function returns 0 or the reference's ERROR_TYPE number where the
// reference would call my_error() and exit; nothing here exits.  Arrays
// the readers return are malloc'd (release them with free()), as the
// reference's are.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/mdgpu.h"

namespace {

constexpr double KCAL = 418.4;                 // kcal/mol -> 10 J/mol (global.h:41)
constexpr double BAR = 6.02214129e-3;          // bar -> internal pressure (global.h:44)
constexpr double RGAS = 0.83144621;            // R in 10 J/(mol K) (global.h:47)
constexpr double WATER_COMPRESS = 0.007372;    // global.h:55
constexpr double PI_D = 3.141592653589793;

enum {
    E_PSF_FILE = 20, E_FROZEN = 21, E_PSF_LINE = 22, E_BOND_SEQ = 23, E_ANGL_SEQ = 24,
    E_DIHE_SEQ = 25, E_IMPR_SEQ = 26, E_CONS_SEQ = 27, E_UB_SEQ = 28, E_FORF_KEY = 29,
    E_CONF_FILE = 40, E_CONF_ATNUM = 41, E_RCON_FILE = 47, E_SIMU_FILE = 60,
    E_SIMU_KEY = 61, E_SIMU_PARAM = 62, E_SIMU_MISSING = 63, E_CONVERG_VEL = 310,
    E_REST_WRITE = 501, E_REST_READ = 502
};

int nint_c(double x) { return x >= 0. ? (int)(x + 0.5) : (int)(x - 0.5); }   // utils.c:606-609

// The whitespace-separated words of one line ( strtok(" \n\t") ).
class Words {
  public:
    explicit Words(const char *line)
    {
        const char *p = line;
        while (*p) {
            while (*p == ' ' || *p == '\n' || *p == '\t') p++;
            if (!*p) break;
            const char *q = p;
            while (*q && *q != ' ' && *q != '\n' && *q != '\t') q++;
            w_.emplace_back(p, q);
            p = q;
        }
    }
    // the next word, or nullptr
    const char *next() { return k_ < w_.size() ? w_[k_++].c_str() : nullptr; }
    // the next word lower-cased, or nullptr
    const char *next_lower()
    {
        if (k_ >= w_.size()) return nullptr;
        std::string &s = w_[k_++];
        for (char &c : s) c = (char)std::tolower((unsigned char)c);
        return s.c_str();
    }

  private:
    std::vector<std::string> w_;
    size_t k_ = 0;
};

struct File {
    FILE *f = nullptr;
    File(const char *name, const char *mode) : f(std::fopen(name, mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
    explicit operator bool() const { return f != nullptr; }
};

template <class T> T *alloc(long n)
{
    return (T *)std::calloc((size_t)(n > 0 ? n : 1), sizeof(T));
}

// ---- SIMU (io.c:29-726) ------------------------------------------------

struct SimuState {
    MDG_IO *io;
    MDG_CTRL *ctrl;
    MDG_PARAM *param;
    MDG_BATH *bath;
    MDG_NEIGH *neigh;
    MDG_EWALD *ewald;
    MDG_PBC *box;
};

void simu_defaults(SimuState &S)
{
    std::strcpy(S.io->confName, "CONF");
    std::strcpy(S.io->forfName, "FORF");
    std::strcpy(S.io->propName, "PROP");
    std::strcpy(S.io->restName, "REST");
    std::strcpy(S.io->rconName, "RCON");
    std::strcpy(S.io->trajName, "TRAJ");
    MDG_CTRL &c = *S.ctrl;
    MDG_PARAM &p = *S.param;
    c.keyMd = 1;
    c.keyRest = 0;
    c.keyMinim = 0;
    p.tolMinim = 1.e-3;
    p.maxminst = 1000;
    p.maxminsiz = 0.15;
    p.step = 0;
    c.newjob = 1;
    c.mdType = 1;
    p.timeStep = 0.001;
    p.nSteps = 0;
    p.cutOff = 12.0;
    p.cutOn = 10.0;
    p.delr = 2.0;
    c.ens = 0;
    p.temp0 = 300.0;
    p.press0 = 1.0;
    S.bath->tauT = 0.1;
    S.bath->tauP = 0.5;
    S.bath->compress = WATER_COMPRESS;
    c.keyRand = 0;
    c.seed = 12345;
    c.keyEwald = 0;
    c.keyAlpha = 0;
    c.keyMmax = 0;
    S.ewald->nbsp = 8;
    c.elecType = 1;   // FULL
    c.vdwType = 1;    // VFULL
    c.keyNb14 = 0;
    p.scal14 = 1.0;
    c.keyNumForce = 0;
    S.neigh->update = 20;
    S.neigh->linkRatio = 1;
    c.noLink = 0;
    c.integrator = 1;   // velocity Verlet
    p.tolShake = 1.e-8;
    p.maxCycle = 150;
    c.keyConstH = 0;
    p.nConst = 0;
    c.keyProp = 0;
    c.keyTraj = 0;
    c.keyForF = 0;
    c.printOut = 1000;
    c.printProp = 1000;
    c.printTraj = 1000;
    c.printRest = 1000;
    S.box->type = 0;
    S.box->a1 = S.box->b2 = S.box->c3 = 0.;
}

// One "name value" option whose value is parsed by the caller.
#define NEED(w) do { if (!(w)) return E_SIMU_MISSING; } while (0)

int simu_file(SimuState &S, Words &W)
{
    const char *kind = W.next_lower();
    NEED(kind);
    char *dst = nullptr;
    if (!std::strcmp(kind, "conf")) dst = S.io->confName;
    else if (!std::strcmp(kind, "forf")) dst = S.io->forfName;
    else if (!std::strcmp(kind, "prop")) dst = S.io->propName;
    else if (!std::strcmp(kind, "rest")) dst = S.io->restName;
    else if (!std::strcmp(kind, "rcon")) dst = S.io->rconName;
    else if (!std::strcmp(kind, "traj")) dst = S.io->trajName;
    else return E_SIMU_PARAM;
    const char *name = W.next();
    NEED(name);
    std::snprintf(dst, MDG_FINAMELEN, "%s", name);
    return 0;
}

int simu_minim(SimuState &S, Words &W)
{
    S.ctrl->keyMinim = 1;
    const char *k = W.next_lower();
    NEED(k);
    const char *v;
    if (!std::strcmp(k, "tol")) {
        NEED(v = W.next());
        S.param->tolMinim = std::atof(v);
    } else if (!std::strcmp(k, "maxcycle")) {
        NEED(v = W.next());
        S.param->maxminst = std::atoi(v);
    } else if (!std::strcmp(k, "maxsize")) {
        NEED(v = W.next());
        S.param->maxminsiz = std::atof(v);
    } else {
        return E_SIMU_PARAM;
    }
    return 0;
}

int simu_ewald(SimuState &S, Words &W)
{
    const char *k = W.next_lower();
    NEED(k);
    const char *v;
    MDG_EWALD &e = *S.ewald;
    if (!std::strcmp(k, "prec")) {
        NEED(v = W.next());
        e.prec = std::atof(v);
    } else if (!std::strcmp(k, "alpha")) {
        NEED(v = W.next());
        S.ctrl->keyAlpha = 1;
        e.alpha = std::atof(v);
    } else if (!std::strcmp(k, "nbsp")) {
        NEED(v = W.next());
        e.nbsp = std::atoi(v);
    } else if (!std::strcmp(k, "recvec")) {
        // read with atof into int fields (io.c:315-327)
        S.ctrl->keyMmax = 1;
        NEED(v = W.next());
        e.m1max = (int)std::atof(v);
        NEED(v = W.next());
        e.m2max = (int)std::atof(v);
        NEED(v = W.next());
        e.m3max = (int)std::atof(v);
    } else {
        return E_SIMU_PARAM;
    }
    return 0;
}

int simu_choice(const char *v, std::initializer_list<std::pair<const char *, int>> table, int *out)
{
    for (const auto &t : table)
        if (!std::strcmp(v, t.first)) {
            *out = t.second;
            return 0;
        }
    return E_SIMU_PARAM;
}

int simu_elec(SimuState &S, Words &W)
{
    const char *v = W.next_lower();
    NEED(v);
    if (!std::strcmp(v, "ewald")) {
        S.ctrl->keyEwald = 1;
        return 0;
    }
    if (!std::strcmp(v, "spme")) {
        S.ctrl->keyEwald = 2;
        return 0;
    }
    return simu_choice(v, {{"noelec", 0}, {"full", 1}, {"shift1", 2}, {"shift2", 3}, {"switch", 4}},
                       &S.ctrl->elecType);
}

int simu_ensemble(SimuState &S, Words &W)
{
    const char *v = W.next_lower();
    NEED(v);
    if (!std::strcmp(v, "nve")) {
        S.ctrl->ens = 0;
        return 0;
    }
    const bool nvt = !std::strcmp(v, "nvt"), npt = !std::strcmp(v, "npt");
    if (!nvt && !npt) return E_SIMU_PARAM;
    const char *b = W.next_lower();
    NEED(b);
    // NVT_B 1, NPT_B 2, NVT_H 3, NPT_H 4 (global.h:121-127)
    if (!std::strcmp(b, "beren")) S.ctrl->ens = nvt ? 1 : 2;
    else if (!std::strcmp(b, "hoover")) S.ctrl->ens = nvt ? 3 : 4;
    else return E_SIMU_PARAM;
    return 0;
}

int simu_shake(SimuState &S, Words &W)
{
    const char *k = W.next_lower();
    NEED(k);
    const char *v;
    if (!std::strcmp(k, "tol")) {
        NEED(v = W.next());
        S.param->tolShake = std::atof(v);
    } else if (!std::strcmp(k, "maxcycle")) {
        NEED(v = W.next());
        S.param->maxCycle = std::atoi(v);
    } else {
        return E_SIMU_PARAM;
    }
    return 0;
}

// The three lines of box vectors after "pbc <type>" (io.c:615-700): every
// number must start with a digit.
int simu_pbc(SimuState &S, Words &W, FILE *f)
{
    const char *v = W.next();
    NEED(v);
    S.box->type = std::atoi(v);
    double *rows[3][3] = {{&S.box->a1, &S.box->a2, &S.box->a3},
                          {&S.box->b1, &S.box->b2, &S.box->b3},
                          {&S.box->c1, &S.box->c2, &S.box->c3}};
    char line[1024];
    for (auto &row : rows) {
        if (!std::fgets(line, sizeof line, f)) return E_SIMU_MISSING;
        Words L(line);
        for (double *d : row) {
            const char *t = L.next();
            NEED(t);
            if (!std::isdigit((unsigned char)t[0])) return E_SIMU_PARAM;
            *d = std::atof(t);
        }
    }
    return 0;
}

int simu_keyword(SimuState &S, const char *key, Words &W, FILE *f, bool *end)
{
    MDG_CTRL &c = *S.ctrl;
    MDG_PARAM &p = *S.param;
    const char *v;
    // run type and flags
    if (!std::strcmp(key, "mdbas")) c.mdType = 0;
    else if (!std::strcmp(key, "charmm")) c.mdType = 1;
    else if (!std::strcmp(key, "namd")) c.mdType = 2;
    else if (!std::strcmp(key, "dlpoly")) c.mdType = 3;
    else if (!std::strcmp(key, "nomd")) c.keyMd = 0;
    else if (!std::strcmp(key, "restart")) c.keyRest = 1;
    else if (!std::strcmp(key, "file")) return simu_file(S, W);
    else if (!std::strcmp(key, "minim")) return simu_minim(S, W);
    else if (!std::strcmp(key, "ewald")) return simu_ewald(S, W);
    else if (!std::strcmp(key, "elec")) return simu_elec(S, W);
    else if (!std::strcmp(key, "vdw")) {
        NEED(v = W.next_lower());
        return simu_choice(v, {{"novdw", 0}, {"full", 1}, {"switch", 2}}, &c.vdwType);
    } else if (!std::strcmp(key, "integrator")) {
        NEED(v = W.next_lower());
        return simu_choice(v, {{"leapfrog", 0}, {"velocity", 1}}, &c.integrator);
    } else if (!std::strcmp(key, "ensemble")) return simu_ensemble(S, W);
    else if (!std::strcmp(key, "shake")) return simu_shake(S, W);
    else if (!std::strcmp(key, "pbc")) return simu_pbc(S, W, f);
    else if (!std::strcmp(key, "numforce")) c.keyNumForce = 1;
    else if (!std::strcmp(key, "nolink")) c.noLink = 1;
    else if (!std::strcmp(key, "consth")) c.keyConstH = 1;
    else if (!std::strcmp(key, "end")) *end = true;
    else {
        // keywords with one value
        static const char *const ONE[] = {"timestep", "nsteps", "cutoff", "cuton", "delr",
                                          "nb14", "list", "link", "taut", "taup",
                                          "temperature", "pressure", "compressibility", "seed",
                                          "print", "prop", "traj", "resconf", "write"};
        bool known = false;
        for (const char *k : ONE) known = known || !std::strcmp(key, k);
        if (!known) return E_SIMU_KEY;
        NEED(v = W.next());
        if (!std::strcmp(key, "timestep")) p.timeStep = std::atof(v);
        else if (!std::strcmp(key, "nsteps")) p.nSteps = std::atoi(v);
        else if (!std::strcmp(key, "cutoff")) p.cutOff = std::atof(v);
        else if (!std::strcmp(key, "cuton")) p.cutOn = std::atof(v);
        else if (!std::strcmp(key, "delr")) p.delr = std::atof(v);
        else if (!std::strcmp(key, "nb14")) {
            c.keyNb14 = 1;
            p.scal14 = std::atof(v);
        } else if (!std::strcmp(key, "list")) S.neigh->update = std::atoi(v);
        else if (!std::strcmp(key, "link")) S.neigh->linkRatio = std::atoi(v);
        else if (!std::strcmp(key, "taut")) S.bath->tauT = std::atof(v);
        else if (!std::strcmp(key, "taup")) S.bath->tauP = std::atof(v);
        else if (!std::strcmp(key, "temperature")) p.temp0 = std::atof(v);
        else if (!std::strcmp(key, "pressure")) p.press0 = std::atof(v) * BAR;
        else if (!std::strcmp(key, "compressibility")) S.bath->compress = std::atof(v) / BAR;
        else if (!std::strcmp(key, "seed")) {
            c.keyRand = 1;
            c.seed = std::atoi(v);
        } else if (!std::strcmp(key, "print")) c.printOut = std::atoi(v);
        else if (!std::strcmp(key, "prop")) {
            c.keyProp = 1;
            c.printProp = std::atoi(v);
        } else if (!std::strcmp(key, "traj")) {
            c.keyTraj = 1;
            c.printTraj = std::atoi(v);
        } else if (!std::strcmp(key, "resconf")) c.printRest = std::atoi(v);
        else if (!std::strcmp(key, "write")) {
            std::string w(v);
            for (char &ch : w) ch = (char)std::tolower((unsigned char)ch);
            if (w != "field") return E_SIMU_PARAM;
            c.keyForF = 1;
        } else {
            return E_SIMU_KEY;
        }
    }
    return 0;
}

#undef NEED

// ---- FORF sections (io.c:872-1199) ------------------------------------

// Reads the next non-comment line of a section into W; false at the end
// of the file.
bool section_line(FILE *f, char *buf, size_t cap)
{
    while (std::fgets(buf, (int)cap, f))
        if (buf[0] != '#') return true;
    return false;
}

struct Num {
    Words W;
    bool ok = true;
    explicit Num(const char *line) : W(line) {}
    int i()
    {
        const char *t = W.next();
        if (!t) ok = false;
        return t ? std::atoi(t) : 0;
    }
    double d()
    {
        const char *t = W.next();
        if (!t) ok = false;
        return t ? std::atof(t) : 0.;
    }
    void skip()
    {
        if (!W.next()) ok = false;
    }
};

}  // namespace

// ======================================================================= readers

extern "C" int mdg_read_simu(MDG_IO *io, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_BATH *bath,
                             MDG_NEIGH *neigh, MDG_EWALD *ewald, MDG_PBC *box)
{
    if (!io || !ctrl || !param || !bath || !neigh || !ewald || !box) return MDG_EARG;
    File f(io->simuName, "r");
    if (!f) return E_SIMU_FILE;
    SimuState S{io, ctrl, param, bath, neigh, ewald, box};
    simu_defaults(S);
    char line[1024];
    bool end = false;
    while (!end && std::fgets(line, sizeof line, f.f)) {
        Words W(line);
        const char *key = W.next_lower();
        if (!key) continue;
        const int st = simu_keyword(S, key, W, f.f, &end);
        if (st) return st;
    }
    // derived cut-off constants (io.c:714-723)
    param->rcutOff = 1. / param->cutOff;
    param->cutOff2 = param->cutOff * param->cutOff;
    param->rcutOff2 = 1. / param->cutOff2;
    param->cutOn2 = param->cutOn * param->cutOn;
    const double d = param->cutOff2 - param->cutOn2;
    param->switch2 = 1. / (d * d * d);
    param->rTimeStep = 1. / param->timeStep;
    return MDG_OK;
}

extern "C" int mdg_read_conf(MDG_IO *io, MDG_PARAM *param, MDG_ATOM **atom, double **x,
                             double **y, double **z)
{
    if (!io || !param || !atom || !x || !y || !z) return MDG_EARG;
    File f(io->confName, "r");
    if (!f) return E_CONF_FILE;
    char line[1024] = "";
    // title lines start with '*'; the first other line holds the count
    while (std::fgets(line, sizeof line, f.f))
        if (line[0] != '*') break;
    Words W(line);
    const char *cnt = W.next();
    if (!cnt) return E_CONF_FILE;
    const int n = std::atoi(cnt);
    if (n <= 0) return E_CONF_ATNUM;
    param->nAtom = n;
    *atom = alloc<MDG_ATOM>(n);
    *x = alloc<double>(n);
    *y = alloc<double>(n);
    *z = alloc<double>(n);
    for (int i = 0; i < n; i++) {
        int serial, ires, resi;
        char resn[5] = "", label[5] = "", segn[5] = "";
        double xx, yy, zz, w;
        // io.c:764 field layout
        if (std::fscanf(f.f, "%d %d %4s %4s %lf %lf %lf %4s %d %lf", &serial, &ires, resn, label,
                        &xx, &yy, &zz, segn, &resi, &w) != 10)
            return E_CONF_ATNUM;
        (*x)[i] = xx;
        (*y)[i] = yy;
        (*z)[i] = zz;
        MDG_ATOM &a = (*atom)[i];
        a.ires = ires;
        a.resi = resi;
        std::strncpy(a.label, label, 4);
        std::strncpy(a.resn, resn, 4);
        std::strncpy(a.segn, segn, 4);
        a.label[4] = a.resn[4] = a.segn[4] = '\0';
    }
    return MDG_OK;
}

extern "C" int mdg_read_rest(MDG_IO *io, MDG_PARAM *param, MDG_ENERGY *ener, MDG_BATH *bath,
                             MDG_ATOM **atom, double **x, double **y, double **z, double **vx,
                             double **vy, double **vz, double **fx, double **fy, double **fz)
{
    if (!io || !param || !ener || !bath || !atom) return MDG_EARG;
    File f(io->restName, "rb");
    if (!f) return E_REST_READ;
    if (std::fread(&param->step, sizeof(int), 1, f.f) != 1) return E_REST_READ;
    if (std::fread(&param->nAtom, sizeof(int), 1, f.f) != 1) return E_REST_READ;
    const int n = param->nAtom;
    if (n <= 0) return E_REST_READ;
    *atom = alloc<MDG_ATOM>(n);
    if (std::fread(*atom, sizeof(MDG_ATOM), n, f.f) != (size_t)n) return E_REST_READ;
    double **arr[9] = {x, y, z, vx, vy, vz, fx, fy, fz};
    for (double **a : arr) {
        *a = alloc<double>(n);
        if (std::fread(*a, sizeof(double), n, f.f) != (size_t)n) return E_REST_READ;
    }
    if (std::fread(ener, sizeof(MDG_ENERGY), 1, f.f) != 1) return E_REST_READ;
    if (std::fread(&bath->chiT, sizeof(double), 1, f.f) != 1) return E_REST_READ;
    if (std::fread(&bath->chiP, sizeof(double), 1, f.f) != 1) return E_REST_READ;
    return MDG_OK;
}

extern "C" int mdg_read_forf(MDG_IO *io, MDG_PARAM *param, MDG_ATOM atom[],
                             MDG_CONSTRAINT **constList, MDG_BOND **bond, MDG_ANGLE **angle,
                             MDG_DIHE **dihe, MDG_DIHE **impr, MDG_BOND **ub, double *eps,
                             double *sig, double *eps14, double *sig14, double *mass, double *q,
                             int *frozen, int *nAtConst)
{
    if (!io || !param || !atom || !constList || !bond || !angle || !dihe || !impr || !ub || !eps ||
        !sig || !eps14 || !sig14 || !mass || !q || !frozen || !nAtConst)
        return MDG_EARG;
    File f(io->forfName, "r");
    if (!f) return E_PSF_FILE;
    param->nBond = param->nConst = param->nUb = param->nAngle = 0;
    param->nDihedral = param->nImproper = 0;
    const int n = param->nAtom;
    char line[1024], sub[1024];
    while (std::fgets(line, sizeof line, f.f)) {
        for (char *p = line; *p; p++) *p = (char)std::tolower((unsigned char)*p);
        if (line[0] == '#') continue;
        Words W(line);
        const char *key = W.next();
        if (!key) continue;
        const char *cntw = W.next();
        const std::string k(key);
        if (k == "end") break;
        if (!cntw) return E_FORF_KEY;
        const int cnt = std::atoi(cntw);
        if (k == "atoms" || k == "vdw") {
            if (cnt != n) return E_CONF_ATNUM;
            for (int r = 0; r < n; r++) {
                if (!section_line(f.f, sub, sizeof sub)) return E_PSF_LINE;
                Num L(sub);
                const int i = L.i() - 1;
                if (i != r) return E_PSF_FILE;
                if (k == "atoms") {
                    L.skip();   // label (CONF's is kept)
                    atom[i].type = L.i();
                    q[i] = L.d();
                    mass[i] = L.d();
                    frozen[i] = L.i();
                    nAtConst[i] = 0;
                    if (!L.ok) return E_PSF_LINE;
                    if (frozen[i] != 0 && frozen[i] != 1) return E_FROZEN;
                } else {
                    L.i();   // type
                    eps[i] = std::sqrt(L.d() * KCAL);
                    sig[i] = L.d();
                    L.d();   // beta
                    eps14[i] = std::sqrt(L.d() * KCAL);
                    sig14[i] = L.d();
                    L.d();
                    if (!L.ok) return E_PSF_LINE;
                }
            }
        } else if (k == "bonds" || k == "urey-bradley") {
            const bool isUb = k == "urey-bradley";
            (isUb ? param->nUb : param->nBond) = cnt;
            if (cnt == 0) continue;
            MDG_BOND *b = alloc<MDG_BOND>(cnt);
            (isUb ? *ub : *bond) = b;
            for (int r = 0; r < cnt; r++) {
                if (!section_line(f.f, sub, sizeof sub)) return isUb ? E_UB_SEQ : E_BOND_SEQ;
                Num L(sub);
                b[r].a = L.i() - 1;
                b[r].b = L.i() - 1;
                b[r].type = isUb ? 0 : L.i();
                b[r].k = L.d() * KCAL;
                b[r].r0 = L.d();
                b[r].beta = L.d();
                if (!L.ok) return isUb ? E_UB_SEQ : E_BOND_SEQ;
            }
        } else if (k == "constraints") {
            param->nConst = cnt;
            if (cnt == 0) continue;
            MDG_CONSTRAINT *c = alloc<MDG_CONSTRAINT>(cnt);
            *constList = c;
            for (int r = 0; r < cnt; r++) {
                if (!section_line(f.f, sub, sizeof sub)) return E_CONS_SEQ;
                Num L(sub);
                c[r].a = L.i() - 1;
                c[r].b = L.i() - 1;
                const double d = L.d();
                c[r].rc2 = d * d;
                if (!L.ok || c[r].a < 0 || c[r].a >= n || c[r].b < 0 || c[r].b >= n)
                    return E_CONS_SEQ;
                nAtConst[c[r].a]++;
                nAtConst[c[r].b]++;
            }
        } else if (k == "angles") {
            param->nAngle = cnt;
            if (cnt == 0) continue;
            MDG_ANGLE *a = alloc<MDG_ANGLE>(cnt);
            *angle = a;
            for (int r = 0; r < cnt; r++) {
                if (!section_line(f.f, sub, sizeof sub)) return E_ANGL_SEQ;
                Num L(sub);
                a[r].a = L.i() - 1;
                a[r].b = L.i() - 1;
                a[r].c = L.i() - 1;
                a[r].type = L.i();
                a[r].k = L.d() * KCAL;
                a[r].theta0 = L.d() * PI_D / 180.;
                if (!L.ok) return E_ANGL_SEQ;
            }
        } else if (k == "dihedrals" || k == "impropers") {
            const bool isImpr = k == "impropers";
            (isImpr ? param->nImproper : param->nDihedral) = cnt;
            if (cnt == 0) continue;
            MDG_DIHE *d = alloc<MDG_DIHE>(cnt);
            (isImpr ? *impr : *dihe) = d;
            for (int r = 0; r < cnt; r++) {
                if (!section_line(f.f, sub, sizeof sub)) return isImpr ? E_IMPR_SEQ : E_DIHE_SEQ;
                Num L(sub);
                d[r].a = L.i() - 1;
                d[r].b = L.i() - 1;
                d[r].c = L.i() - 1;
                d[r].d = L.i() - 1;
                d[r].type = L.i();
                d[r].order = L.i();
                d[r].k = L.d() * KCAL;
                d[r].phi0 = L.d() * PI_D / 180.;
                d[r].mult = L.d();
                if (!L.ok) return isImpr ? E_IMPR_SEQ : E_DIHE_SEQ;
            }
        } else {
            return E_FORF_KEY;
        }
    }
    return MDG_OK;
}

// ======================================================================= set-up

extern "C" double mdg_charge_const(int mdType)
{
    // init.c:102-109: CHARMM, NAMD, DL_POLY, else the SI value
    switch (mdType) {
    case 1: return 332.0716 * KCAL;
    case 2: return 332.0636 * KCAL;
    case 3: return 138935.4835;
    default: {
        const double mu0 = 1.e-7, c = 299792458., e = 1.602176565e-19, na = 6.02214129e+23;
        return mu0 * (c * c) * (e * e) * na * 0.1 / 1.e-10;
    }
    }
}

// init.c:840-908
extern "C" void mdg_init_box(MDG_PBC *b)
{
    b->a = std::sqrt(b->a1 * b->a1 + b->a2 * b->a2 + b->a3 * b->a3);
    b->b = std::sqrt(b->b1 * b->b1 + b->b2 * b->b2 + b->b3 * b->b3);
    b->c = std::sqrt(b->c1 * b->c1 + b->c2 * b->c2 + b->c3 * b->c3);
    // reciprocal (cofactor) vectors
    b->u1 = b->b2 * b->c3 - b->c2 * b->b3;
    b->u2 = b->c1 * b->b3 - b->b1 * b->c3;
    b->u3 = b->b1 * b->c2 - b->c1 * b->b2;
    b->v1 = b->c2 * b->a3 - b->a2 * b->c3;
    b->v2 = b->a1 * b->c3 - b->c1 * b->a3;
    b->v3 = b->c1 * b->a2 - b->a1 * b->c2;
    b->w1 = b->a2 * b->b3 - b->b2 * b->a3;
    b->w2 = b->b1 * b->a3 - b->a1 * b->b3;
    b->w3 = b->a1 * b->b2 - b->b1 * b->a2;
    b->det = b->a1 * b->u1 + b->a2 * b->u2 + b->a3 * b->u3;
    b->vol = std::fabs(b->det);
    b->vol0 = b->vol;
    // widths of the slab between opposite faces
    b->pa = b->vol / std::sqrt(b->u1 * b->u1 + b->u2 * b->u2 + b->u3 * b->u3);
    b->pb = b->vol / std::sqrt(b->v1 * b->v1 + b->v2 * b->v2 + b->v3 * b->v3);
    b->pc = b->vol / std::sqrt(b->w1 * b->w1 + b->w2 * b->w2 + b->w3 * b->w3);
    b->u = 1.0 / b->pa;
    b->v = 1.0 / b->pb;
    b->w = 1.0 / b->pc;
    double *r[9] = {&b->u1, &b->u2, &b->u3, &b->v1, &b->v2, &b->v3, &b->w1, &b->w2, &b->w3};
    const bool ok = std::fabs(b->det) > 0.;
    for (double *p : r) *p = ok ? *p / b->det : 0.0;
}

// utils.c:569-582 then 551-561
extern "C" void mdg_get_kinfromtemp(MDG_PARAM *p, const MDG_PBC *box)
{
    p->nDegFree = 3 * (p->nAtom - p->nFrozen) - 3;
    if (box->type == 0) p->nDegFree -= 3;
    if (p->nConst > 0) p->nDegFree -= p->nConst;
    p->kinTemp0 = 0.5 * p->temp0 * (double)p->nDegFree * RGAS;
}

// utils.c:162-218, in place
extern "C" void mdg_image_update(const MDG_PARAM *p, const MDG_PBC *b, double x[], double y[],
                                 double z[])
{
    const int n = p->nAtom;
    switch (b->type) {
    case 1:
        for (int i = 0; i < n; i++) {
            x[i] = x[i] - b->a1 * nint_c(x[i] * b->u1);
            y[i] = y[i] - b->a1 * nint_c(y[i] * b->u1);
            z[i] = z[i] - b->a1 * nint_c(z[i] * b->u1);
        }
        break;
    case 2:
        for (int i = 0; i < n; i++) {
            x[i] = x[i] - b->a1 * nint_c(x[i] * b->u1);
            y[i] = y[i] - b->b2 * nint_c(y[i] * b->v2);
            z[i] = z[i] - b->c3 * nint_c(z[i] * b->w3);
        }
        break;
    case 3:
        for (int i = 0; i < n; i++) {
            double s1 = x[i] * b->u1 + y[i] * b->u2 + z[i] * b->u3;
            double s2 = x[i] * b->v1 + y[i] * b->v2 + z[i] * b->v3;
            double s3 = x[i] * b->w1 + y[i] * b->w2 + z[i] * b->w3;
            s1 -= nint_c(s1);
            s2 -= nint_c(s2);
            s3 -= nint_c(s3);
            x[i] = s1 * b->a1 + s2 * b->b1 + s3 * b->c1;
            y[i] = s1 * b->a2 + s2 * b->b2 + s3 * b->c2;
            z[i] = s1 * b->a3 + s2 * b->b3 + s3 * b->c3;
        }
        break;
    default: break;
    }
}

// spme.c:32-73: Ewald width (unless given), the tolerances and the mesh
// (each edge doubled from the next power of two of the requested count,
// at most 512); writes ewald->alpha, tol, tol1, m1max..m3max, mmax.
extern "C" int mdg_spme_mesh(const MDG_CTRL *ctrl, const MDG_PARAM *p, MDG_EWALD *e,
                             const MDG_PBC *b)
{
    if (!ctrl || !p || !e || !b) return MDG_EARG;
    e->prec = std::fmin(std::fabs(e->prec), 0.5);
    e->tol = std::sqrt(std::fabs(std::log(e->prec * p->cutOff)));
    if (!ctrl->keyAlpha)
        e->alpha = std::sqrt(std::fabs(std::log(e->prec * p->cutOff * e->tol))) / p->cutOff;
    const double t2 = 2.0 * e->tol * e->alpha;
    e->tol1 = std::sqrt(-std::log(e->prec * p->cutOff * (t2 * t2)));
    if (!ctrl->keyMmax) {
        e->m1max = nint_c(0.25 + b->pa * e->alpha * e->tol1 / PI_D);
        e->m2max = nint_c(0.25 + b->pb * e->alpha * e->tol1 / PI_D);
        e->m3max = nint_c(0.25 + b->pc * e->alpha * e->tol1 / PI_D);
    }
    int *m[3] = {&e->m1max, &e->m2max, &e->m3max};
    for (int *mk : m) {
        int p2 = 1;
        while (*mk > p2 && p2 < 256) p2 *= 2;
        *mk = 2 * p2;
    }
    e->mmax = e->m1max * e->m2max * e->m3max;
    return MDG_OK;
}

// init.c:248-616: frozen atoms' masses, constraints between two frozen
// atoms dropped, bonds involving a hydrogen turned into constraints under
// consth, Urey-Bradley terms and angles spanned entirely by constraints
// dropped, terms among frozen atoms dropped; the term arrays are shrunk
// (released when empty).  Keeps the reference's constraint counts as they
// are: a constraint read from the FORF "constraints" section that is kept
// lowers the per-atom count that read_FORF raised (init.c:286-291).
extern "C" int mdg_setup(MDG_CTRL *ctrl, MDG_PARAM *p, MDG_ATOM atom[],
                         MDG_CONSTRAINT **constList, MDG_BOND **bond, MDG_ANGLE **angle,
                         MDG_DIHE **dihe, MDG_DIHE **impr, MDG_BOND **ub, double mass[],
                         double rmass[], int frozen[], int nAtConst[])
{
    if (!ctrl || !p || !atom || !constList || !bond || !angle || !dihe || !impr || !ub || !mass ||
        !rmass || !frozen || !nAtConst)
        return MDG_EARG;
    const int n = p->nAtom;
    p->nFrozen = 0;
    for (int i = 0; i < n; i++) {
        p->nFrozen += frozen[i];
        if (frozen[i] || mass[i] < 2.220446049250313e-16) {
            mass[i] = 0.;
            rmass[i] = 0.;
        } else {
            rmass[i] = 1.0 / mass[i];
        }
    }
    auto both = [&](int a, int b) { return frozen[a] && frozen[b]; };
    // constraints of the FORF section
    std::vector<MDG_CONSTRAINT> cons;
    for (int i = 0; i < p->nConst; i++) {
        const MDG_CONSTRAINT c = (*constList)[i];
        if (!both(c.a, c.b)) {
            cons.push_back(c);
            nAtConst[c.a]--;
            nAtConst[c.b]--;
        }
    }
    // bonds: under consth, a bond with a hydrogen becomes a constraint
    std::vector<MDG_BOND> bonds;
    for (int i = 0; i < p->nBond; i++) {
        const MDG_BOND b = (*bond)[i];
        if (both(b.a, b.b)) continue;
        if (ctrl->keyConstH && (atom[b.a].label[0] == 'H' || atom[b.b].label[0] == 'H')) {
            cons.push_back(MDG_CONSTRAINT{b.a, b.b, b.r0 * b.r0});
            nAtConst[b.a]++;
            nAtConst[b.b]++;
        } else {
            bonds.push_back(b);
        }
    }
    auto constrained = [&](int a, int b) {
        for (const MDG_CONSTRAINT &c : cons)
            if ((c.a == a && c.b == b) || (c.a == b && c.b == a)) return true;
        return false;
    };
    std::vector<MDG_BOND> ubs;
    for (int i = 0; i < p->nUb; i++) {
        const MDG_BOND b = (*ub)[i];
        if (both(b.a, b.b)) continue;
        if (ctrl->keyConstH && constrained(b.a, b.b)) continue;
        ubs.push_back(b);
    }
    std::vector<MDG_ANGLE> angles;
    for (int i = 0; i < p->nAngle; i++) {
        const MDG_ANGLE a = (*angle)[i];
        if (frozen[a.a] && frozen[a.b] && frozen[a.c]) continue;
        if (ctrl->keyConstH && constrained(a.a, a.b) && constrained(a.a, a.c) &&
            constrained(a.c, a.b))
            continue;
        angles.push_back(a);
    }
    auto keep4 = [&](MDG_DIHE **arr, int cnt) {
        std::vector<MDG_DIHE> out;
        for (int i = 0; i < cnt; i++) {
            const MDG_DIHE d = (*arr)[i];
            if (!(frozen[d.a] && frozen[d.b] && frozen[d.c] && frozen[d.d])) out.push_back(d);
        }
        return out;
    };
    std::vector<MDG_DIHE> dihes = keep4(dihe, p->nDihedral), imprs = keep4(impr, p->nImproper);
    // shrink (or release) every array to its kept terms
    auto store = [](auto **arr, const auto &v, int *count) {
        using T = typename std::decay<decltype(v[0])>::type;
        std::free(*arr);
        *arr = nullptr;
        *count = (int)v.size();
        if (!v.empty()) {
            *arr = (T *)std::malloc(v.size() * sizeof(T));
            std::memcpy(*arr, v.data(), v.size() * sizeof(T));
        }
    };
    store(constList, cons, &p->nConst);
    store(bond, bonds, &p->nBond);
    store(ub, ubs, &p->nUb);
    store(angle, angles, &p->nAngle);
    store(dihe, dihes, &p->nDihedral);
    store(impr, imprs, &p->nImproper);
    return MDG_OK;
}

// utils.c:502-515
extern "C" double mdg_kinetic(const MDG_PARAM *p, const double vx[], const double vy[],
                              const double vz[], const double mass[])
{
    double e = 0.;
    for (int i = 0; i < p->nAtom; i++)
        e += mass[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    return 0.5 * e;
}

// init.c:627-833 with the Gaussian deviates given: gauss holds 3N
// standard normal numbers (vx of every atom, then vy, then vz), in place of
// the reference's Box-Muller calls on its dSFMT stream.  Weighted by
// 1/sqrt(m), projected onto the constraints (init_constvel: every
// constraint's relative velocity along its bond removed, iterated to
// tolShake within maxCycle sweeps, else 310), centre-of-mass motion
// removed, scaled to kinTemp0.
extern "C" int mdg_init_vel(const MDG_PARAM *p, const MDG_PBC *b, const MDG_CONSTRAINT cons[],
                            const double x[], const double y[], const double z[], double vx[],
                            double vy[], double vz[], const double mass[], const double rmass[],
                            const int frozen[], const int nAtConst[], const double gauss[])
{
    if (!p || !b || !x || !vx || !mass || !rmass || !frozen || !gauss) return MDG_EARG;
    const int n = p->nAtom;
    for (int i = 0; i < n; i++) {
        const double s = std::sqrt(rmass[i]);
        vx[i] = gauss[i] * s;
        vy[i] = gauss[n + i] * s;
        vz[i] = gauss[2 * n + i] * s;
    }
    const int nc = p->nConst;
    if (nc > 0) {
        if (!cons || !nAtConst) return MDG_EARG;
        std::vector<double> ux(nc), uy(nc), uz(nc), dx(n), dy(n), dz(n);
        for (int k = 0; k < nc; k++) {
            const int a = cons[k].a, c = cons[k].b;
            double ex = x[c] - x[a], ey = y[c] - y[a], ez = z[c] - z[a];
            // minimum image (utils.c:82-122)
            switch (b->type) {
            case 1:
                ex -= b->a1 * nint_c(ex * b->u1);
                ey -= b->a1 * nint_c(ey * b->u1);
                ez -= b->a1 * nint_c(ez * b->u1);
                break;
            case 2:
                ex -= b->a1 * nint_c(ex * b->u1);
                ey -= b->b2 * nint_c(ey * b->v2);
                ez -= b->c3 * nint_c(ez * b->w3);
                break;
            case 3: {
                double s1 = ex * b->u1 + ey * b->u2 + ez * b->u3;
                double s2 = ex * b->v1 + ey * b->v2 + ez * b->v3;
                double s3 = ex * b->w1 + ey * b->w2 + ez * b->w3;
                s1 -= nint_c(s1);
                s2 -= nint_c(s2);
                s3 -= nint_c(s3);
                ex = s1 * b->a1 + s2 * b->b1 + s3 * b->c1;
                ey = s1 * b->a2 + s2 * b->b2 + s3 * b->c2;
                ez = s1 * b->a3 + s2 * b->b3 + s3 * b->c3;
                break;
            }
            default: break;
            }
            const double r = std::sqrt((ex * ex + ey * ey) + ez * ez);
            ux[k] = ex / r;
            uy[k] = ey / r;
            uz[k] = ez / r;
        }
        bool converged = false;
        for (int cyc = 0; cyc < p->maxCycle && !converged; cyc++) {
            std::fill(dx.begin(), dx.end(), 0.);
            std::fill(dy.begin(), dy.end(), 0.);
            std::fill(dz.begin(), dz.end(), 0.);
            double worst = 0.;
            for (int k = 0; k < nc; k++) {
                const int a = cons[k].a, c = cons[k].b;
                const double dv = ux[k] * (vx[c] - vx[a]) + uy[k] * (vy[c] - vy[a]) +
                                  uz[k] * (vz[c] - vz[a]);
                worst = std::fmax(worst, std::fabs(dv));
                const double wa = mass[c] * dv / (mass[a] + mass[c]);
                const double wc = mass[a] * dv / (mass[a] + mass[c]);
                dx[a] += wa * ux[k];
                dy[a] += wa * uy[k];
                dz[a] += wa * uz[k];
                dx[c] -= wc * ux[k];
                dy[c] -= wc * uy[k];
                dz[c] -= wc * uz[k];
            }
            converged = worst < p->tolShake;
            if (converged) break;
            // each constraint applies its atoms' summed change / count
            for (int k = 0; k < nc; k++) {
                const int a = cons[k].a, c = cons[k].b;
                const double na = (double)nAtConst[a], ncn = (double)nAtConst[c];
                vx[a] += dx[a] / na;
                vy[a] += dy[a] / na;
                vz[a] += dz[a] / na;
                vx[c] += dx[c] / ncn;
                vy[c] += dy[c] / ncn;
                vz[c] += dz[c] / ncn;
            }
        }
        if (!converged) return E_CONVERG_VEL;
    }
    double px = 0., py = 0., pz = 0., m = 0.;
    const double eps = 2.220446049250313e-16;
    for (int i = 0; i < n; i++)
        if (!frozen[i] && mass[i] >= eps) {
            px += vx[i] * mass[i];
            py += vy[i] * mass[i];
            pz += vz[i] * mass[i];
            m += mass[i];
        }
    px /= m;
    py /= m;
    pz /= m;
    for (int i = 0; i < n; i++) {
        const bool free_ = !frozen[i] && mass[i] >= eps;
        vx[i] = free_ ? vx[i] - px : 0.;
        vy[i] = free_ ? vy[i] - py : 0.;
        vz[i] = free_ ? vz[i] - pz : 0.;
    }
    const double f = std::sqrt(p->kinTemp0 / mdg_kinetic(p, vx, vy, vz, mass));
    for (int i = 0; i < n; i++) {
        vx[i] *= f;
        vy[i] *= f;
        vz[i] *= f;
    }
    return MDG_OK;
}

// ======================================================================= writers

extern "C" int mdg_write_conf(MDG_IO *io, const MDG_PARAM *p, const MDG_ATOM atom[],
                              const double x[], const double y[], const double z[])
{
    if (!io || !p || !atom || !x || !y || !z) return MDG_EARG;
    File f(io->rconName, "w");
    if (!f) return E_RCON_FILE;
    std::fprintf(f.f, "*Restart Configuration\n* Written by MDBas\n*\n%5d\n", p->nAtom);
    for (int i = 0; i < p->nAtom; i++)
        std::fprintf(f.f, "%5d%5d %-4s %-4s%10.5lf%10.5lf%10.5lf %-4s %-4d%10.5lf\n", i + 1,
                     atom[i].ires, atom[i].resn, atom[i].label, x[i], y[i], z[i], atom[i].segn,
                     atom[i].resi, 0.);
    return MDG_OK;
}

// One PROP record: 29 doubles appended (io.c:1230-1274).
extern "C" int mdg_write_prop(MDG_IO *io, const MDG_PARAM *p, const MDG_ENERGY *e,
                              const MDG_PBC *b)
{
    if (!io || !p || !e || !b) return MDG_EARG;
    double r[29] = {0.};
    r[0] = (double)p->step;
    r[1] = (double)p->step * p->timeStep;
    r[2] = 2. * e->kin / ((double)p->nDegFree * RGAS);
    r[3] = b->type > 0 ? (2. * e->kin - e->virtot) / (3. * b->vol * BAR) : 0.;
    r[4] = b->vol;
    const double v[18] = {e->tot, e->kin, e->pot, e->elec, e->vdw, e->bond, e->ang, e->ub,
                          e->dihe, e->impr, e->virtot, e->virpot, e->virelec, e->virvdw,
                          e->virbond, e->virub, e->virshake, e->consv};
    for (int k = 0; k < 18; k++) r[5 + k] = v[k] / KCAL;
    // cell lengths and angles (utils.c:456-473)
    r[23] = b->a;
    r[24] = b->b;
    r[25] = b->c;
    r[26] = std::acos((b->b1 * b->c1 + b->b2 * b->c2 + b->b3 * b->c3) / (b->b * b->c)) * 180. / PI_D;
    r[27] = std::acos((b->a1 * b->c1 + b->a2 * b->c2 + b->a3 * b->c3) / (b->a * b->c)) * 180. / PI_D;
    r[28] = std::acos((b->a1 * b->b1 + b->a2 * b->b2 + b->a3 * b->b3) / (b->a * b->b)) * 180. / PI_D;
    File f(io->propName, "ab");
    if (!f) return E_REST_WRITE;
    if (std::fwrite(r, sizeof(double), 29, f.f) != 29) return E_REST_WRITE;
    return MDG_OK;
}

extern "C" int mdg_write_rest(MDG_IO *io, const MDG_PARAM *p, const MDG_ENERGY *e,
                              const MDG_BATH *bath, const MDG_ATOM atom[], const double x[],
                              const double y[], const double z[], const double vx[],
                              const double vy[], const double vz[], const double fx[],
                              const double fy[], const double fz[])
{
    if (!io || !p || !e || !bath || !atom) return MDG_EARG;
    File f(io->restName, "wb");
    if (!f) return E_REST_WRITE;
    const size_t n = (size_t)p->nAtom;
    bool ok = std::fwrite(&p->step, sizeof(int), 1, f.f) == 1 &&
              std::fwrite(&p->nAtom, sizeof(int), 1, f.f) == 1 &&
              std::fwrite(atom, sizeof(MDG_ATOM), n, f.f) == n;
    const double *arr[9] = {x, y, z, vx, vy, vz, fx, fy, fz};
    for (const double *a : arr) ok = ok && a && std::fwrite(a, sizeof(double), n, f.f) == n;
    ok = ok && std::fwrite(e, sizeof(MDG_ENERGY), 1, f.f) == 1 &&
         std::fwrite(&bath->chiT, sizeof(double), 1, f.f) == 1 &&
         std::fwrite(&bath->chiP, sizeof(double), 1, f.f) == 1;
    return ok ? MDG_OK : E_REST_WRITE;
}

namespace {

// Fortran-style unformatted record: length, payload, length.
bool record(FILE *f, const void *data, unsigned int bytes)
{
    return std::fwrite(&bytes, sizeof bytes, 1, f) == 1 &&
           (bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes) &&
           std::fwrite(&bytes, sizeof bytes, 1, f) == 1;
}

long dcd_header_bytes(const MDG_PARAM *p)
{
    long h = (8 + 84) + (8 + 4 + 240) + (8 + 4);
    if (p->nFrozen > 0) h += 8 + 4L * (p->nAtom - p->nFrozen);
    return h;
}

}  // namespace

// The CHARMM DCD header (io.c:1371-1460): "CORD" and the control words,
// three 80-character title lines (date, user and host, working directory),
// the atom count and, with frozen atoms, the indices of the free ones.
extern "C" int mdg_write_dcd_header(MDG_IO *io, const MDG_CTRL *ctrl, const MDG_PARAM *p,
                                    const MDG_PBC *b, const int frozen[])
{
    if (!io || !ctrl || !p || !b || !frozen || ctrl->printTraj <= 0) return MDG_EARG;
    File f(io->trajName, "wb");
    if (!f) return E_REST_WRITE;
    struct {
        char hdr[4];
        int icntrl[20];
    } h1 = {{'C', 'O', 'R', 'D'}, {0}};
    h1.icntrl[0] = p->nSteps / ctrl->printTraj;
    h1.icntrl[1] = ctrl->printTraj;
    h1.icntrl[2] = ctrl->printTraj;
    h1.icntrl[3] = p->nSteps;
    h1.icntrl[7] = p->nDegFree;
    h1.icntrl[8] = p->nFrozen;
    h1.icntrl[10] = b->type == 0 ? 0 : 1;
    h1.icntrl[19] = 37;
    struct {
        int ntitle;
        char title[3][80];
    } h2;
    std::memset(&h2, 0, sizeof h2);
    h2.ntitle = 3;
    const time_t now = std::time(nullptr);
    const char *user = std::getenv("USER"), *host = std::getenv("HOSTNAME"),
               *pwd = std::getenv("PWD");
    char tmp[4096];
    std::snprintf(tmp, sizeof tmp, "* CREATION DATE : %s", std::asctime(std::localtime(&now)));
    std::strncat(h2.title[0], tmp, 79);
    std::snprintf(tmp, sizeof tmp, "* USER : %s HOSTNAME : %s", user ? user : "UNKNOWN",
                  host ? host : "UNKNOWN");
    std::strncat(h2.title[1], tmp, 79);
    std::snprintf(tmp, sizeof tmp, "* PWD : %s", pwd ? pwd : "UNKNOWN");
    std::strncat(h2.title[2], tmp, 79);
    const int natom = p->nAtom;
    bool ok = record(f.f, &h1, sizeof h1) && record(f.f, &h2, sizeof h2) &&
              record(f.f, &natom, sizeof natom);
    if (ok && p->nFrozen > 0) {
        std::vector<int> freeAt;
        for (int i = 0; i < natom; i++)
            if (!frozen[i]) freeAt.push_back(i);
        ok = record(f.f, freeAt.data(), (unsigned)(freeAt.size() * sizeof(int)));
    }
    return ok ? MDG_OK : E_REST_WRITE;
}

// One DCD frame (io.c:1462-1522): the box's lower triangle for periodic
// boxes, then x, y, z as float records.  The first frame after the header
// holds every atom; later ones only the free atoms.  The reference tells
// the first frame by a static flag; here it is the frame written to a file
// that holds only the header.
extern "C" int mdg_write_dcd_traj(MDG_IO *io, const MDG_PARAM *p, const MDG_PBC *b,
                                  const double x[], const double y[], const double z[],
                                  const int frozen[])
{
    if (!io || !p || !b || !x || !y || !z || !frozen) return MDG_EARG;
    File f(io->trajName, "ab");
    if (!f) return E_REST_WRITE;
    std::fseek(f.f, 0, SEEK_END);
    const bool first = std::ftell(f.f) == dcd_header_bytes(p);
    bool ok = true;
    if (b->type != 0) {
        const double cr[6] = {b->a1, b->a2, b->b2, b->a3, b->b3, b->c3};   // utils.c:475-492
        ok = record(f.f, cr, sizeof cr);
    }
    std::vector<float> X, Y, Z;
    for (int i = 0; i < p->nAtom; i++) {
        if (frozen[i] && !first) continue;
        X.push_back((float)x[i]);
        Y.push_back((float)y[i]);
        Z.push_back((float)z[i]);
    }
    const unsigned bytes = (unsigned)(X.size() * sizeof(float));
    ok = ok && record(f.f, X.data(), bytes) && record(f.f, Y.data(), bytes) &&
         record(f.f, Z.data(), bytes);
    return ok ? MDG_OK : E_REST_WRITE;
}
