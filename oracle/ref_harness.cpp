// ref_harness.cpp -- drives the real xDEM reference (compiled from its own
// sources by oracle/Makefile into oracle/_ref/) for golden vectors and for
// the CPU baseline timing.  TEST INFRASTRUCTURE ONLY.
//
// Modes
//   ref_harness run   <state.xst> <out.xdg> <nsteps> <dump_every> [threads]
//       builds the reference world from a state file, runs takeTimeStep and
//       dumps particle state, forces, contacts, stress, LE scalars and the
//       world's observables.
//   ref_harness pairs <pairs.bin> <out.bin>
//       calls EllipticParticle / CircularParticle::calculateContactGeometry
//       on a list of poses (src/EllipticParticle.cpp:16-32).
//   ref_harness bench <state.xst> <warmup> <steps> <threads> <chunk>
//       wall-clock timing of takeTimeStep, prints one JSON line.
//   ref_harness control <state.xst> <out.xdg> <nsteps> <mode> <stress> <rate_x>
//                       <rate_y> <rattler_scale> <reset_rattlers> <mass> <max_rate>
//       a PeriodicControlWorld run in the order of the drivers' compression
//       loop (mainIsoElpsPack_WCCM.cpp:127-137): scaleVelocityOfRattlers, the
//       step, resetVelocityOfRattlers, updateTotalStress, boundary update
//       (mode 1 strain, 2 stress uniform, 3 stress X then Y, 4 stress uniform
//       with strain limit); dumps every step plus the boundary record.
//   ref_harness history <state.xst> <prefix> <nsteps> <every>
//       the reference's own file writers: writeParticleInformation2Files and
//       createParticleTimeHistoryFiles(prefix), then every `every` steps
//       updateTotalStress + writeParticleTimeHistory2Files.  The reference
//       joins names with '\\', so the files are "<prefix>\\output_*.dat".
//   ref_harness checkpoint <state.xst> <ckpt.txt> <restart.xdg> <k> <m>
//       k steps, then the text checkpoint BaseWorld::print2File_allParticles
//       (src/BaseWorld.cpp:369-375); a second world is built from that file
//       with the reference's stream constructor EllipticParticle(std::ifstream&,
//       aspectRatio, contactModel, ifReadContact = true) (src/EllipticParticle.h:40-48,
//       read4File_particle src/BaseParticle.cpp:173-186), the boundary and
//       Lees-Edwards shift of the first, and runs m steps, dumped after each.
//       Ellipses only (the circle's stream constructor sets no contact model).
//   ref_harness quasistatic <state.xst> <out.xdg> <nsteps>
//       the compression loop of mainIsoElpsPackTestEntropy.cpp:140-180 in a
//       PeriodicControlWorld: the step's parts, scaleVelocityOfRattlers(0.95),
//       the observables, and updatePeriodicBoundary_strainControl(-0.2, -0.2)
//       whenever the driver's criterion holds; dumps every step, then the
//       boundary record and 1.0 when the boundary moved (magic XDEMGQ01).
//
// The state file layout is written by xdem_amd/state.py (write_state).
#include <chrono>
#include <cstdio>
#include <fstream>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include <algorithm>

#include "BaseParticle.h"
#include "EllipticParticle.h"
#include "CircularParticle.h"
#include "GluedParticle.h"
#include "PeriodicBoundaryWorld.h"
#include "PeriodicControlWorld.h"
#include "LeesEdwardsBoundaryWorld.h"

namespace {

struct State {
    int n = 0, world = 0;
    double cfg[14];   // xmin,xmax,ymin,ymax,dt,kn,kt,e,mu,gdamp,rate,shift,strain,cell
    std::vector<int> kind, up, flags;
    std::vector<double> r, ar, m, inertia, x, y, th, vx, vy, w, cbx, cby;
    std::vector<int> ci, cj;
    std::vector<double> xix, xiy, nx, ny, c;
};

template <class T> void rd(FILE* f, T* p, size_t n) {
    if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(3); }
}

State read_state(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(3); }
    char magic[8];
    rd(f, magic, 8);
    if (memcmp(magic, "XDEMST02", 8) != 0) { fprintf(stderr, "bad magic\n"); exit(3); }
    State s;
    int hdr[2];
    rd(f, hdr, 2);
    s.n = hdr[0]; s.world = hdr[1];
    rd(f, s.cfg, 14);
    auto ri = [&](std::vector<int>& v) { v.resize(s.n); rd(f, v.data(), s.n); };
    auto rdv = [&](std::vector<double>& v) { v.resize(s.n); rd(f, v.data(), s.n); };
    ri(s.kind); ri(s.up); ri(s.flags);
    rdv(s.r); rdv(s.ar); rdv(s.m); rdv(s.inertia); rdv(s.x); rdv(s.y); rdv(s.th);
    rdv(s.vx); rdv(s.vy); rdv(s.w); rdv(s.cbx); rdv(s.cby);
    int nc = 0;
    rd(f, &nc, 1);
    s.ci.resize(nc); s.cj.resize(nc);
    rd(f, s.ci.data(), nc); rd(f, s.cj.data(), nc);
    for (auto* v : {&s.xix, &s.xiy, &s.nx, &s.ny, &s.c}) { v->resize(nc); rd(f, v->data(), nc); }
    fclose(f);
    return s;
}

// Subclasses only to reach protected members of the reference classes.
struct HEllipse : EllipticParticle {
    using EllipticParticle::EllipticParticle;
    void inject(int j, double xix, double xiy, double nx, double ny, double c) {
        ContactInformation ci(j, vector2d(nx, ny), vector2d(0, 0), vector2d(0, 0), vector2d(0, 0));
        ci._shearOverlap = vector2d(xix, xiy);
        ci._dampingConstant = c;
        _contactInfo[j] = ci;
    }
};
struct HCircle : CircularParticle {
    using CircularParticle::CircularParticle;
    void inject(int j, double xix, double xiy, double nx, double ny, double c) {
        ContactInformation ci(j, vector2d(nx, ny), vector2d(0, 0), vector2d(0, 0), vector2d(0, 0));
        ci._shearOverlap = vector2d(xix, xiy);
        ci._dampingConstant = c;
        _contactInfo[j] = ci;
    }
};
struct HGlued : GluedParticle {
    HGlued() {}
    HGlued(int id, vector2d v, double w) : GluedParticle(id, v, w) {}
    void setup(double x, double y, double th, double m, double I) {
        _position = vector2d(x, y); _rotation = th;
        _mass = m; _invMass = 1.0 / m; _inertia = I; _invInertia = 1.0 / I;
        _volume = 0.0;
    }
    void add_branch(double bx, double by) { _branchVectors.push_back(vector2d(bx, by)); }
};
template <class Base> struct HWorld : Base {
    HWorld(double dt, const vector<BaseParticle*>& p, ContactModel* cm) : Base(dt, p, cm) {}
    const vector<vector2d>& force() const { return this->_totalForce; }
    const vector<double>& torque() const { return this->_totalTorque; }
    const vector<int>& ncont() const { return this->_numContactPerParticle; }
};
struct HLE : HWorld<LeesEdwardsBoundaryWorld> {
    using HWorld::HWorld;
    void set_shift(double s, double e) { _totalShift = s; _totalStrain = e; }
    void scalars(double* o) const { o[0] = _totalShift; o[1] = _shiftVel; o[2] = _totalStrain; }
};
using HPB = HWorld<PeriodicBoundaryWorld>;
struct HPC : HWorld<PeriodicControlWorld> {
    using HWorld::HWorld;
    // xmin, xmax, ymin, ymax, boundary velocity x, y, total stress
    void box(double* o) const {
        o[0] = _xmin; o[1] = _xmax; o[2] = _ymin; o[3] = _ymax; o[4] = _velX; o[5] = _velY;
        o[6] = _totalStress.xx(); o[7] = _totalStress.xy();
        o[8] = _totalStress.yx(); o[9] = _totalStress.yy();
    }
};

struct Built {
    State s;
    std::unique_ptr<ContactModel> cm;
    std::vector<std::unique_ptr<BaseParticle>> parts;
    vector<BaseParticle*> handles;
    std::unique_ptr<HLE> le;
    std::unique_ptr<HPB> pb;
    std::unique_ptr<HPC> pc;
    BaseWorld* world() {
        return le ? (BaseWorld*)le.get() : pc ? (BaseWorld*)pc.get() : (BaseWorld*)pb.get();
    }
};

void make_world(Built& B, int threads, int chunk, bool control);

void build(Built& B, int threads, int chunk, bool control = false) {
    State& s = B.s;
    const double* c = s.cfg;
    B.cm.reset(new ContactModel(c[5], c[6], c[7], c[8]));
    BaseParticle::_GlobalDamping_ = c[9];
    for (int i = 0; i < s.n; ++i) {
        vector2d p(s.x[i], s.y[i]), v(s.vx[i], s.vy[i]);
        BaseParticle* bp;
        if (s.kind[i] == 0) {
            auto* q = new HCircle(i, s.r[i], p, v, s.th[i], s.w[i], B.cm.get());
            q->calculateParticleProperties(1.0);
            bp = q;
        } else if (s.kind[i] == 1) {
            auto* q = new HEllipse(i, s.r[i], s.ar[i], p, v, s.th[i], s.w[i], B.cm.get());
            q->calculateParticleProperties(1.0);
            bp = q;
        } else {
            auto* q = new HGlued(i, v, s.w[i]);
            q->setup(s.x[i], s.y[i], s.th[i], s.m[i], s.inertia[i]);
            bp = q;
        }
        if (s.kind[i] != 2) { bp->setMass(s.m[i]); bp->setInertia(s.inertia[i]); }
        bp->setUpLevel(s.up[i]);
        if (s.flags[i]) bp->appendSetting((SETTING)s.flags[i]);
        B.parts.emplace_back(bp);
        B.handles.push_back(bp);
    }
    for (int i = 0; i < s.n; ++i) {       // clump membership
        if (s.kind[i] != 2) continue;
        auto* g = (HGlued*)B.parts[i].get();
        for (int k = i + 1; k < s.n && s.up[k] == i; ++k) {
            g->pushBackSubParticle(B.parts[k].get());
            g->add_branch(s.cbx[k], s.cby[k]);
        }
    }
    for (size_t k = 0; k < s.ci.size(); ++k) {
        BaseParticle* p = B.parts[s.ci[k]].get();
        if (s.kind[s.ci[k]] == 0)
            ((HCircle*)p)->inject(s.cj[k], s.xix[k], s.xiy[k], s.nx[k], s.ny[k], s.c[k]);
        else
            ((HEllipse*)p)->inject(s.cj[k], s.xix[k], s.xiy[k], s.nx[k], s.ny[k], s.c[k]);
    }
    make_world(B, threads, chunk, control);
}

void make_world(Built& B, int threads, int chunk, bool control) {
    const State& s = B.s;
    const double* c = s.cfg;
    if (s.world == 1) {
        B.le.reset(new HLE(c[4], B.handles, B.cm.get()));
        B.le->setWorldBoundary(c[0], c[1], c[2], c[3]);
        B.le->setShiftStrainRate(c[10]);
        B.le->set_shift(c[11], c[12]);
    } else if (control) {
        B.pc.reset(new HPC(c[4], B.handles, B.cm.get()));
        B.pc->setWorldBoundary(c[0], c[1], c[2], c[3]);
    } else {
        B.pb.reset(new HPB(c[4], B.handles, B.cm.get()));
        B.pb->setWorldBoundary(c[0], c[1], c[2], c[3]);
    }
    B.world()->set_OMP_NumThreadsAndSizeChunk(threads, chunk);
    B.world()->updateTotalParticleVolume();
}

template <class T> void wr(FILE* f, const T* p, size_t n) { if (n) fwrite(p, sizeof(T), n, f); }

void dump(Built& B, FILE* f, int step) {
    const int n = B.s.n;
    BaseWorld* W = B.world();
    std::vector<double> a(n);
    int hdr[2] = {step, n};
    wr(f, hdr, 2);
    for (int k = 0; k < 6; ++k) {
        for (int i = 0; i < n; ++i) {
            BaseParticle* p = B.handles[i];
            a[i] = k == 0 ? p->getPosX() : k == 1 ? p->getPosY() : k == 2 ? p->getRot()
                 : k == 3 ? p->getVelX() : k == 4 ? p->getVelY() : p->getRotVel();
        }
        wr(f, a.data(), n);
    }
    const vector<vector2d>* F; const vector<double>* T; const vector<int>* NC;
    if (B.le) { F = &B.le->force(); T = &B.le->torque(); NC = &B.le->ncont(); }
    else if (B.pc) { F = &B.pc->force(); T = &B.pc->torque(); NC = &B.pc->ncont(); }
    else { F = &B.pb->force(); T = &B.pb->torque(); NC = &B.pb->ncont(); }
    for (int i = 0; i < n; ++i) a[i] = (*F)[i].x();
    wr(f, a.data(), n);
    for (int i = 0; i < n; ++i) a[i] = (*F)[i].y();
    wr(f, a.data(), n);
    wr(f, T->data(), n);
    wr(f, NC->data(), n);
    // contacts ordered by (i, j)
    std::vector<int> ci, cj;
    std::vector<double> cd;
    for (int i = 0; i < n; ++i) {
        auto m = B.handles[i]->getContactInfo();
        std::vector<std::pair<int, const ContactInformation*>> v;
        for (auto& kv : m) v.push_back({kv.first, &kv.second});
        std::sort(v.begin(), v.end(), [](auto& x, auto& y) { return x.first < y.first; });
        for (auto& e : v) {
            const ContactInformation& c = *e.second;
            ci.push_back(i); cj.push_back(e.first);
            double rec[13] = {c._normalForce.x(), c._normalForce.y(), c._shearForce.x(),
                              c._shearForce.y(), c._contactPoint.x(), c._contactPoint.y(),
                              c._shearOverlap.x(), c._shearOverlap.y(), c._contactNormal.x(),
                              c._contactNormal.y(), c._dampingConstant, c._torqueOnP1,
                              c._torqueOnP2};
            cd.insert(cd.end(), rec, rec + 13);
        }
    }
    int nc = (int)ci.size();
    wr(f, &nc, 1);
    wr(f, ci.data(), nc);
    wr(f, cj.data(), nc);
    wr(f, cd.data(), cd.size());
    matrix2d st = W->updateTotalStress();
    double sv[4] = {st.xx(), st.xy(), st.yx(), st.yy()};
    wr(f, sv, 4);
    double sc[3] = {0, 0, 0};
    if (B.le) B.le->scalars(sc);
    wr(f, sc, 3);
    // observables (src/BaseWorld.cpp:251-334)
    double ob[8] = {(double)W->countPossibleContacts(), (double)W->countActucalContacts(),
                    W->getKineticEnergy(), W->getKineticEnergyPerParticle(),
                    W->getKineticEnergyPerNonRattlerParticle(), W->getElasticEnergy(),
                    W->getElasticEnergyPerContact(), W->getSolidFraction()};
    wr(f, ob, 8);
}

int mode_run(int argc, char** argv) {
    if (argc < 6) return 2;
    Built B;
    B.s = read_state(argv[2]);
    int nsteps = atoi(argv[4]), every = atoi(argv[5]);
    int threads = argc > 6 ? atoi(argv[6]) : 1;
    build(B, threads, 100);
    FILE* f = fopen(argv[3], "wb");
    if (!f) return 3;
    fwrite("XDEMGD02", 1, 8, f);
    for (int s = 1; s <= nsteps; ++s) {
        B.world()->takeTimeStep();
        if (every > 0 && (s % every == 0 || s == nsteps)) dump(B, f, s);
    }
    fclose(f);
    return 0;
}

// pairs file: int32 m; then m records of 12 doubles
//   kind, r1, ar1, x1, y1, th1, r2, ar2, x2, y2, th2, (unused)
// out: m records of 10 doubles: contact, overlap, px, py, nx, ny, b1x, b1y, b2x, b2y
int mode_pairs(int argc, char** argv) {
    if (argc < 4) return 2;
    FILE* f = fopen(argv[2], "rb");
    if (!f) return 3;
    int m = 0;
    rd(f, &m, 1);
    std::vector<double> in((size_t)m * 12);
    rd(f, in.data(), in.size());
    fclose(f);
    ContactModel cm(1e4, 5e3, 0.5, 0.5);
    std::vector<double> out((size_t)m * 10, 0.0);
    for (int k = 0; k < m; ++k) {
        const double* q = &in[(size_t)k * 12];
        ContactGeometry g;
        bool hit;
        if ((int)q[0] == 0) {
            CircularParticle a(0, q[1], vector2d(q[3], q[4]), vector2d(0, 0), q[5], 0, &cm);
            CircularParticle b(1, q[6], vector2d(q[8], q[9]), vector2d(0, 0), q[10], 0, &cm);
            hit = a.calculateContactGeometry(&b, g);
        } else {
            EllipticParticle a(0, q[1], q[2], vector2d(q[3], q[4]), vector2d(0, 0), q[5], 0, &cm);
            EllipticParticle b(1, q[6], q[7], vector2d(q[8], q[9]), vector2d(0, 0), q[10], 0, &cm);
            hit = a.calculateContactGeometry(&b, g);
        }
        double* o = &out[(size_t)k * 10];
        o[0] = hit ? 1.0 : 0.0;
        if (hit) {
            o[1] = g.normalOverlap; o[2] = g.contactPoint.x(); o[3] = g.contactPoint.y();
            o[4] = g.contactNormal.x(); o[5] = g.contactNormal.y();
            o[6] = g.branchVectorFromP1.x(); o[7] = g.branchVectorFromP1.y();
            o[8] = g.branchVectorFromP2.x(); o[9] = g.branchVectorFromP2.y();
        }
    }
    f = fopen(argv[3], "wb");
    if (!f) return 3;
    wr(f, out.data(), out.size());
    fclose(f);
    return 0;
}

int mode_bench(int argc, char** argv) {
    if (argc < 7) return 2;
    Built B;
    B.s = read_state(argv[2]);
    int warm = atoi(argv[3]), steps = atoi(argv[4]), threads = atoi(argv[5]), chunk = atoi(argv[6]);
    build(B, threads, chunk);
    for (int s = 0; s < warm; ++s) B.world()->takeTimeStep();
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) B.world()->takeTimeStep();
    auto t1 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>(t1 - t0).count();
    printf("{\"n\": %d, \"steps\": %d, \"threads\": %d, \"seconds\": %.6f, "
           "\"particle_steps_per_s\": %.6e}\n",
           B.s.n, steps, threads, sec, (double)B.s.n * steps / sec);
    return 0;
}

int mode_history(int argc, char** argv) {
    if (argc < 6) return 2;
    Built B;
    B.s = read_state(argv[2]);
    const std::string prefix = argv[3];
    const int nsteps = atoi(argv[4]), every = atoi(argv[5]);
    build(B, 1, 100);
    BaseWorld* W = B.world();
    W->writeParticleInformation2Files(prefix);
    W->createParticleTimeHistoryFiles(prefix);
    for (int s = 1; s <= nsteps; ++s) {
        W->takeTimeStep();
        if (s % every == 0) {
            W->updateTotalStress();
            W->writeParticleTimeHistory2Files();
        }
    }
    W->flushAllFiles();
    W->closeParticleTimeHistoryFiles();
    return 0;
}

int mode_checkpoint(int argc, char** argv) {
    if (argc < 7) return 2;
    Built B;
    B.s = read_state(argv[2]);
    const int k = atoi(argv[5]), m = atoi(argv[6]);
    for (int i = 0; i < B.s.n; ++i)
        if (B.s.kind[i] != 1) { fprintf(stderr, "checkpoint: ellipses only\n"); return 2; }
    build(B, 1, 100);
    for (int s = 1; s <= k; ++s) B.world()->takeTimeStep();
    FILE* t = fopen(argv[3], "w");
    if (!t) return 3;
    B.world()->print2File_allParticles(t);
    fclose(t);
    double sc[3] = {0, 0, 0};
    if (B.le) B.le->scalars(sc);
    Built R;
    R.s = B.s;
    const double* c = R.s.cfg;
    R.cm.reset(new ContactModel(c[5], c[6], c[7], c[8]));
    std::ifstream is(argv[3]);
    int n = 0;
    is >> n;
    if (n != R.s.n) return 3;
    for (int i = 0; i < n; ++i) {
        auto* q = new HEllipse(is, R.s.ar[i], R.cm.get(), true);
        R.parts.emplace_back(q);
        R.handles.push_back(q);
    }
    make_world(R, 1, 100, false);
    if (R.le) R.le->set_shift(sc[0], sc[2]);
    FILE* f = fopen(argv[4], "wb");
    if (!f) return 3;
    fwrite("XDEMGD02", 1, 8, f);
    for (int s = 1; s <= m; ++s) {
        R.world()->takeTimeStep();
        dump(R, f, s);
    }
    fclose(f);
    return 0;
}

int mode_quasistatic(int argc, char** argv) {
    if (argc < 5) return 2;
    Built B;
    B.s = read_state(argv[2]);
    const int nsteps = atoi(argv[4]);
    build(B, 1, 100, true);
    HPC& W = *B.pc;
    W.updateTotalParticleVolume();
    W.setVirtualBoundaryMass();
    FILE* f = fopen(argv[3], "wb");
    if (!f) return 3;
    fwrite("XDEMGQ01", 1, 8, f);
    int count = 0, count0 = 0;
    bool goNextComp = true;
    const int critCntNum = B.s.n * 2.5;
    for (int s = 1; s <= nsteps; ++s) {
        W.modifyParticlePosition();
        W.findPossibleContacts();
        W.updateContacts();
        W.collectForceAndTorque();
        W.takeTimeIntegral();
        W.scaleVelocityOfRattlers(0.95);
        const double elasticEnergy = W.getElasticEnergyPerContact();
        const double kineticEnergy = W.getKineticEnergyPerNonRattlerParticle();
        const int actualContactNum = W.countActucalContacts();
        double moved = 0.0;
        if ((elasticEnergy < 1e-6 && actualContactNum < critCntNum && goNextComp) ||
            (elasticEnergy >= 1e-6 && actualContactNum >= critCntNum && kineticEnergy < 5e-2 &&
             goNextComp)) {
            if (count == 1) {
                W.updateTotalStress();
                count = 0;
                count0 = 0;
            }
            W.updatePeriodicBoundary_strainControl(-2e-1, -2e-1);
            count++;
            goNextComp = false;
            moved = 1.0;
        } else {
            count0++;
            if (count0 > 10) goNextComp = true;
        }
        double bx[10];
        W.box(bx);
        dump(B, f, s);
        wr(f, bx, 10);
        wr(f, &moved, 1);
    }
    fclose(f);
    return 0;
}

int mode_control(int argc, char** argv) {
    if (argc < 13) return 2;
    Built B;
    B.s = read_state(argv[2]);
    const int nsteps = atoi(argv[4]), mode = atoi(argv[5]);
    const double stress = atof(argv[6]), rx = atof(argv[7]), ry = atof(argv[8]);
    const double rscale = atof(argv[9]);
    const int reset = atoi(argv[10]);
    const double mass = atof(argv[11]), maxrate = atof(argv[12]);
    build(B, 1, 100, true);
    HPC& W = *B.pc;
    W.updateTotalParticleVolume();
    if (mass > 0) W.setVirtualBoundaryMass(mass);
    else W.setVirtualBoundaryMass();
    W.setMaxStrainRate(maxrate);
    FILE* f = fopen(argv[3], "wb");
    if (!f) return 3;
    fwrite("XDEMGC01", 1, 8, f);
    for (int s = 1; s <= nsteps; ++s) {
        if (rscale != 1.0) W.scaleVelocityOfRattlers(rscale);
        W.modifyParticlePosition();
        W.findPossibleContacts();
        W.updateContacts();
        W.collectForceAndTorque();
        W.takeTimeIntegral();
        if (reset) W.resetVelocityOfRattlers();
        W.updateTotalStress();
        if (mode == 1) W.updatePeriodicBoundary_strainControl(rx, ry);
        else if (mode == 2) W.updatePeriodicBoundary_stressControlUniform(stress);
        else if (mode == 3) {
            W.updatePeriodicBoundary_stressControlX(stress);
            W.updatePeriodicBoundary_stressControlY(stress);
        } else if (mode == 4) W.updatePeriodicBoundary_stressControlUniform_strainLimit(stress);
        double bx[10];
        W.box(bx);
        dump(B, f, s);
        wr(f, bx, 10);
    }
    fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) { fprintf(stderr, "usage: ref_harness run|pairs|bench ...\n"); return 2; }
    // the reference prints "Logger msg" lines; keep stdout for JSON only
    std::string m = argv[1];
    if (m == "run") return mode_run(argc, argv);
    if (m == "pairs") return mode_pairs(argc, argv);
    if (m == "bench") return mode_bench(argc, argv);
    if (m == "control") return mode_control(argc, argv);
    if (m == "history") return mode_history(argc, argv);
    if (m == "checkpoint") return mode_checkpoint(argc, argv);
    if (m == "quasistatic") return mode_quasistatic(argc, argv);
    return 2;
}
