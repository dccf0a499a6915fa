// lps_liphen_adapter.hpp -- the HIP backend as a drop-in for the reference's
// solver classes (GitDaroth/LiquidSimulation, LiquidPhysics/).
//
// The reference runs one step as PhysicSolver::update(dt)
// (LiquidPhysics/src/PhysicSolver.cpp:12-21): onBeginUpdate, accumulateForces,
// integrate, handleCollisions, onEndUpdate, each branching on
// ParallelizationType {NONE, CPU, GPU} (LiquidPhysics/include/SPHSolver.h:19-23;
// SPHSolver.cpp:186-199, 384-429).  LiPhEn::HipSolver<Base> derives from
// LiPhEn::SPHSolver or LiPhEn::PCISPHSolver and, while the HIP backend is
// enabled, routes those phases to liblps (include/lps.h):
//
//   onBeginUpdate     uploads what changed since the last step: parameters
//                     (compared with the last ones sent), colliders
//                     (m_hasCollisionObjectDataChanged, boxes then spheres as
//                     SPHSolver.cpp:932-1003 packs them) and particles
//                     (m_hasParticleDataChanged, SPHSolver.cpp:791);
//   accumulateForces  runs the whole step on the device (lps_step);
//   integrate, handleCollisions
//                     nothing (done by that step);
//   onEndUpdate       writes positions, velocities, half velocities and
//                     first-step flags back into the particle objects in the
//                     step's sorted order, as SPHSolver.cpp:412-422 does
//                     (skipped when headless).
//
// With the backend disabled, or when no gfx950 device is found (the
// reference's silent fallback, SPHSolver.cpp:74-85), every phase is the base
// class's.  A viewer selects it where it constructs its solvers
// (LiquidSimulation.h:70-71) and maps its fourth combo-box entry
// (LiquidSimulation.cpp:454-464) to setHipEnabled(true).
//
// Needs one change in the reference: SPHSolver::m_collisionObjects
// (SPHSolver.h, private section) must be protected so that the colliders can
// be read here; the viewer adds them through SPHSolver::addStaticCollisionObject,
// which is not virtual.
#pragma once

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "lps.h"

#include "Collision/StaticCollisionBox.h"
#include "Collision/StaticCollisionSphere.h"
#include "PCISPHSolver.h"
#include "SPHSolver.h"

namespace LiPhEn {

template <class Base>
class HipSolver : public Base {
   public:
    static constexpr bool kPCISPH = std::is_base_of<PCISPHSolver, Base>::value;

    HipSolver() = default;
    ~HipSolver() override {
        if (m_lps) lps_destroy(m_lps);
    }

    // The fourth backend: on (HIP) or off (the base class's NONE/CPU/GPU).
    void setHipEnabled(bool on) {
        m_hip = on;
        m_sent = false;
        this->m_hasParticleDataChanged = true;
        this->m_hasCollisionObjectDataChanged = true;
    }
    bool isHipEnabled() const { return m_hip && m_lps != nullptr; }
    // Skip the per-step readback into the particle objects (benchmarks).
    void setHeadless(bool headless) { m_headless = headless; }
    // lps_status of the last device step (the reference has no error channel).
    int getLastStatus() const { return m_status; }
    const lps_step_stats& getLastStepStats() const { return m_stats; }
    // PCISPH knobs the reference lacks (include/lps.h).
    void setMaxIterations(int n) { m_maxIterations = n; }
    void setStallIterations(int n) { m_stallIterations = n; }
    void setWarmStart(int mode) { m_warmStart = mode; }
    lps_ctx* context() { return m_lps; }

   protected:
    void onBeginUpdate() override {
        if (!m_hip) return Base::onBeginUpdate();
        const lps_params p = params();
        if (!m_lps) {
            if (lps_create(&p, &m_lps) != LPS_OK) {
                m_lps = nullptr;
                m_hip = false;  // no device: the reference's path, silently
                return Base::onBeginUpdate();
            }
            m_last = p;
            m_sent = true;
        } else if (!m_sent || std::memcmp(&p, &m_last, sizeof(p)) != 0) {
            m_status = lps_set_params(m_lps, &p);
            m_last = p;
            m_sent = true;
        }
        if (this->m_hasCollisionObjectDataChanged) {
            std::vector<lps_box> boxes;
            std::vector<lps_sphere> spheres;
            for (StaticCollisionObject* o : this->m_collisionObjects) {
                const bool boundary = o->getType() == StaticCollisionObjectType::BOUNDARY;
                const Vector3D c = o->getPosition();
                if (const auto* b = dynamic_cast<const StaticCollisionBox*>(o)) {
                    const Vector3D h = b->getHalfDimensions();
                    boxes.push_back(lps_box{{c.getX(), c.getY(), c.getZ()}, {h.getX(), h.getY(), h.getZ()},
                                            boundary ? 1 : 0});
                } else if (const auto* s = dynamic_cast<const StaticCollisionSphere*>(o)) {
                    spheres.push_back(lps_sphere{{c.getX(), c.getY(), c.getZ()}, s->getRadius(), boundary ? 1 : 0});
                }
            }
            m_status = lps_set_colliders(m_lps, boxes.data(), (uint32_t)boxes.size(), spheres.data(),
                                         (uint32_t)spheres.size(), nullptr, 0);
            this->m_hasCollisionObjectDataChanged = false;
        }
        if (this->m_hasParticleDataChanged) {
            const size_t n = this->m_particles.size();
            std::vector<float> x(3 * n), v(3 * n), vh(3 * n);
            std::vector<uint8_t> first(n);
            for (size_t i = 0; i < n; i++) {
                const SPHParticle* q = this->m_particles[i];
                put(x, i, q->getPosition());
                put(v, i, q->getVelocity());
                put(vh, i, q->getHalfVelocity());
                first[i] = q->getIsFirstTimeStep() ? 1 : 0;
            }
            m_status = n ? lps_set_particles(m_lps, x.data(), v.data(), vh.data(), first.data(), (uint32_t)n)
                         : LPS_OK;
            this->m_hasParticleDataChanged = false;
        }
    }

    // SPHSolver::accumulateForces is private and virtual; overriding it here
    // is where the step's time step is known.  Without the backend it does
    // what the base's does (SPHSolver.cpp:307-311), which it cannot call.
    void accumulateForces(float deltaTime) override {
        if (!isHipEnabled()) {
            this->accumulateNonPressureForces(deltaTime);
            this->accumulatePressureForces(deltaTime);
            return;
        }
        if (this->m_particles.empty()) return;
        m_status = lps_step(m_lps, deltaTime, 1, &m_stats);
    }

    void integrate(float deltaTime) override {
        if (!isHipEnabled()) Base::integrate(deltaTime);
    }

    void handleCollisions() override {
        if (!isHipEnabled()) Base::handleCollisions();
    }

    void onEndUpdate() override {
        if (!isHipEnabled()) return Base::onEndUpdate();
        const size_t n = this->m_particles.size();
        if (m_headless || n == 0 || m_status != LPS_OK) return;
        std::vector<float> x(3 * n), v(3 * n), vh(3 * n);
        std::vector<uint8_t> first(n);
        m_status = lps_get_particles(m_lps, x.data(), v.data(), vh.data(), first.data(), nullptr, (uint32_t)n);
        if (m_status != LPS_OK) return;
        for (size_t i = 0; i < n; i++) {
            SPHParticle* q = this->m_particles[i];
            q->setPosition(get(x, i));
            q->setVelocity(get(v, i));
            q->setHalfVelocity(get(vh, i));
            q->setIsFirstTimeStep(first[i] != 0);
        }
    }

   private:
    static void put(std::vector<float>& a, size_t i, const Vector3D& u) {
        a[3 * i] = u.getX();
        a[3 * i + 1] = u.getY();
        a[3 * i + 2] = u.getZ();
    }
    static Vector3D get(const std::vector<float>& a, size_t i) {
        return Vector3D(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
    }

    // The solver's parameters as lps_params (field for field the reference's
    // setters; PCISPH's m_minIterations is the reference NONE loop's cap with
    // an early exit, PCISPHSolver.cpp:64,158-159: min 1, max m_minIterations).
    lps_params params() const {
        lps_params p;
        std::memset(&p, 0, sizeof(p));
        p.abi_version = LPS_ABI_VERSION;
        p.method = kPCISPH ? LPS_METHOD_PCISPH : LPS_METHOD_SPH;
        const Vector3D g = this->getGravity();
        p.gravity[0] = g.getX();
        p.gravity[1] = g.getY();
        p.gravity[2] = g.getZ();
        p.particle_radius = this->getParticleRadius();
        p.kernel_radius_factor = this->getKernelRadiusFactor();
        p.rest_density = this->getRestDensity();
        p.stiffness = this->getPressureStiffnessCoefficient();
        p.negative_pressure_factor = this->getNegativePressureFactor();
        p.viscosity = this->getViscosityCoefficient();
        p.surface_tension = this->getSurfaceTensionCoefficient();
        p.surface_tension_threshold = this->getSurfaceTensionThreshold();
        p.restitution = this->getResitutionCoefficient();
        p.friction = this->getFrictionCoefficient();
        p.mass_override = 0.f;
        p.min_iterations = 1;
        p.max_iterations = 1;
        p.max_density_error_ratio = 0.f;
        if constexpr (kPCISPH) {
            p.max_iterations = m_maxIterations > 0 ? m_maxIterations : this->getMinIterations();
            p.max_density_error_ratio = this->getMaxDensityErrorRatio();
        }
        p.warm_start = m_warmStart;
        p.stall_iterations = m_stallIterations;
        return p;
    }

    lps_ctx* m_lps = nullptr;
    bool m_hip = false;
    bool m_sent = false;
    bool m_headless = false;
    int m_status = LPS_OK;
    int m_maxIterations = 0;    // 0: the reference's m_minIterations
    int m_stallIterations = 0;  // the reference has no stall rule
    int m_warmStart = LPS_WARM_NONE;
    lps_params m_last{};
    lps_step_stats m_stats{};
};

using HipSPHSolver = HipSolver<SPHSolver>;
using HipPCISPHSolver = HipSolver<PCISPHSolver>;

}  // namespace LiPhEn
