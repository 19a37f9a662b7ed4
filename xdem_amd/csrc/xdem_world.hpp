// xdem_world.hpp -- header-only C++ adapter over the C-ABI of libxdem_hip.so
// that keeps the method names of the reference's worlds, so a driver loop
// written against xDEM's PeriodicBoundaryWorld / LeesEdwardsBoundaryWorld
// (src/BaseWorld.h:44-117, src/LeesEdwardsBoundaryWorld.h:15-56) compiles
// with little change:
//
//     xdem::ContactModelParams cm{1e4, 5e3, 0.5, 0.5};      // ContactModel(kn, kt, e, mu)
//     std::vector<xdem::ParticleSpec> parts = ...;           // EllipticParticle(...) data
//     xdem::GpuLeesEdwardsBoundaryWorld world(dt, parts, cm);
//     world.setWorldBoundary(xmin, xmax, ymin, ymax);
//     world.setShiftStrainRate(1e-4);
//     for (int s = 0; s < n; ++s) world.takeTimeStep();
//     matrix2 sigma = world.updateTotalStress();
//
// Errors: the reference prints and calls exit(); this adapter throws
// xdem::Error carrying the XDEM_E_* code and the library's message.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xdem.h"

namespace xdem {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ContactModel_LinearDissipativeFrictioanl(kn, kt, restitution, friction),
// src/ContactModel_LinearDissipativeFrictional.h:14-23
struct ContactModelParams {
    double kn, kt, restitution, mu;
};

// One particle as the reference's constructors take it
// (src/EllipticParticle.h:19-30, src/CircularParticle.h:18-27), plus the
// results of calculateParticleProperties and the SETTING flags.
struct ParticleSpec {
    int kind = XDEM_KIND_ELLIPSE;       // XDEM_KIND_*
    double outerRadius = 0.5, aspectRatio = 1.0;
    double x = 0, y = 0, vx = 0, vy = 0, rotation = 0, rotationVel = 0;
    double mass = 0, inertia = 0;       // 0: computed from density 1
    int upLevel = -1;                   // -1: own ID
    int settings = 0;                   // XDEM_SKIP_* bits
    double clumpBx = 0, clumpBy = 0;    // sub-particle branch from its clump
};

struct matrix2 {
    double xx, xy, yx, yy;
    double trace() const { return xx + yy; }
};

// calculateParticleProperties with density 1 (src/EllipticParticle.cpp:7-14,
// src/CircularParticle.cpp:16-23)
inline void fillProperties(ParticleSpec& p) {
    if (p.mass > 0) return;
    const double r = p.outerRadius;
    if (p.kind == XDEM_KIND_CIRCLE) {
        p.mass = M_PI * r * r;
        p.inertia = p.mass * r * r / 2.0;
    } else {
        p.mass = M_PI * r * r / p.aspectRatio;
        p.inertia = p.mass * (r * r + r * r / p.aspectRatio / p.aspectRatio) / 4.0;
    }
}

class GpuWorld {
public:
    GpuWorld(int worldKind, double dt, const std::vector<ParticleSpec>& particles,
             const ContactModelParams& cm)
        : parts_(particles) {
        cfg_.world_kind = worldKind;
        cfg_.precision = XDEM_PREC_FP64;
        cfg_.dt = dt;
        cfg_.kn = cm.kn; cfg_.kt = cm.kt; cfg_.restitution = cm.restitution; cfg_.mu = cm.mu;
        for (auto& p : parts_) fillProperties(p);
    }
    virtual ~GpuWorld() { if (e_) xdem_destroy(e_); }
    GpuWorld(const GpuWorld&) = delete;
    GpuWorld& operator=(const GpuWorld&) = delete;

    // -- set-up (BaseWorld.h:91-101) ---------------------------------------
    void setWorldBoundary(double xmin, double xmax, double ymin, double ymax) {
        cfg_.xmin = xmin; cfg_.xmax = xmax; cfg_.ymin = ymin; cfg_.ymax = ymax;
    }
    void setTimeIncrement(double dt) { cfg_.dt = dt; }
    void setGlobalDamping(double g) { cfg_.global_damping = g; }
    void set_OMP_NumThreadsAndSizeChunk(int, int) {}   // no host threads to size
    void setPrecision(int precision) { cfg_.precision = precision; }

    // -- the step (BaseWorld.h:59-71) --------------------------------------
    void takeTimeStep() { takeTimeSteps(1); }
    void takeTimeSteps(int n) {
        upload();
        check(xdem_step(e_, n));
        dirty_ = true;
    }

    // -- observables -------------------------------------------------------
    const matrix2 updateTotalStress() {        // src/BaseWorld.cpp:158-176
        upload();
        double s[4];
        check(xdem_get_stress(e_, s));
        stress_ = {s[0], s[1], s[2], s[3]};
        return stress_;
    }
    const matrix2 getStress() const { return stress_; }
    const int countActucalContacts() {         // src/BaseWorld.cpp:262-269
        upload();
        int nc = 0;
        check(xdem_get_contacts(e_, 0, &nc, nullptr, nullptr, nullptr, nullptr, nullptr,
                                nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                nullptr, nullptr, nullptr, nullptr));
        return nc;
    }
    const double getKineticEnergy() {          // src/BaseWorld.cpp:302-310
        sync_state();
        double k = 0.0;
        for (const auto& p : parts_)
            k += 0.5 * (p.mass * (p.vx * p.vx + p.vy * p.vy) + p.inertia * p.rotationVel * p.rotationVel);
        return k;
    }
    const double getSolidFraction() const {    // src/BaseWorld.cpp:331-334
        double v = 0.0;
        for (const auto& p : parts_)
            if (!(p.settings & XDEM_SKIP_MACRO_SUM))
                v += p.kind == XDEM_KIND_CIRCLE ? M_PI * p.outerRadius * p.outerRadius
                                                : M_PI * p.outerRadius * p.outerRadius / p.aspectRatio;
        return v / ((cfg_.xmax - cfg_.xmin) * (cfg_.ymax - cfg_.ymin));
    }
    // particle state in ID order (BaseParticle getters)
    const std::vector<ParticleSpec>& particles() { sync_state(); return parts_; }
    void getForces(std::vector<double>& fx, std::vector<double>& fy, std::vector<double>& tq) {
        upload();
        fx.resize(parts_.size()); fy.resize(parts_.size()); tq.resize(parts_.size());
        check(xdem_get_forces(e_, fx.data(), fy.data(), tq.data(), nullptr));
    }

protected:
    void upload() {
        if (e_) return;
        check(xdem_create(&e_, 1, &cfg_));
        const int n = (int)parts_.size();
        std::vector<int> kind(n), up(n), fl(n);
        std::vector<double> r(n), ar(n), m(n), I(n), x(n), y(n), th(n), vx(n), vy(n), w(n), bx(n), by(n);
        for (int i = 0; i < n; ++i) {
            const auto& p = parts_[i];
            kind[i] = p.kind; up[i] = p.upLevel < 0 ? i : p.upLevel; fl[i] = p.settings;
            r[i] = p.outerRadius; ar[i] = p.aspectRatio; m[i] = p.mass; I[i] = p.inertia;
            x[i] = p.x; y[i] = p.y; th[i] = p.rotation; vx[i] = p.vx; vy[i] = p.vy;
            w[i] = p.rotationVel; bx[i] = p.clumpBx; by[i] = p.clumpBy;
        }
        check(xdem_set_particles(e_, n, nullptr, kind.data(), up.data(), fl.data(), r.data(),
                                 ar.data(), m.data(), I.data(), x.data(), y.data(), th.data(),
                                 vx.data(), vy.data(), w.data(), bx.data(), by.data()));
    }
    void sync_state() {
        if (!e_ || !dirty_) return;
        const size_t n = parts_.size();
        std::vector<double> x(n), y(n), th(n), vx(n), vy(n), w(n);
        check(xdem_get_particles(e_, x.data(), y.data(), th.data(), vx.data(), vy.data(), w.data()));
        for (size_t i = 0; i < n; ++i) {
            auto& p = parts_[i];
            p.x = x[i]; p.y = y[i]; p.rotation = th[i]; p.vx = vx[i]; p.vy = vy[i]; p.rotationVel = w[i];
        }
        dirty_ = false;
    }
    void check(int rc) const {
        if (rc != XDEM_OK) throw Error(rc, xdem_last_error(e_));
    }

    xdem_config cfg_{};
    xdem_engine* e_ = nullptr;
    std::vector<ParticleSpec> parts_;
    matrix2 stress_{0, 0, 0, 0};
    bool dirty_ = false;
};

// PeriodicBoundaryWorld(dt, particleHandlers, contactModel), src/PeriodicBoundaryWorld.h:13-19
class GpuPeriodicBoundaryWorld : public GpuWorld {
public:
    GpuPeriodicBoundaryWorld(double dt, const std::vector<ParticleSpec>& p, const ContactModelParams& cm)
        : GpuWorld(XDEM_WORLD_PERIODIC, dt, p, cm) {}
};

// LeesEdwardsBoundaryWorld(dt, particleHandlers, contactModel), src/LeesEdwardsBoundaryWorld.h:18-26
class GpuLeesEdwardsBoundaryWorld : public GpuWorld {
public:
    GpuLeesEdwardsBoundaryWorld(double dt, const std::vector<ParticleSpec>& p, const ContactModelParams& cm)
        : GpuWorld(XDEM_WORLD_LEES_EDWARDS, dt, p, cm) {}
    // setShiftStrainRate, src/LeesEdwardsBoundaryWorld.cpp:485-489 (after setWorldBoundary)
    void setShiftStrainRate(double rate) { cfg_.shear_rate = rate; }
    const double getTotalShift() { return scalars()[0]; }
    const double getTotalStrain() { return scalars()[2]; }

private:
    std::vector<double> scalars() {
        upload();
        std::vector<double> b(3);
        check(xdem_get_boundary(e_, b.data()));
        return b;
    }
};

}  // namespace xdem
