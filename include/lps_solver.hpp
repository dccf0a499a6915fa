/*
 * include/lps_solver.hpp -- C++ host layer over the C ABI (include/lps.h):
 * the public interface of the reference's solvers, stepping on the device.
 *
 *   lps::SPHSolver     LiPhEn::SPHSolver     LiquidPhysics/include/SPHSolver.h:25-69
 *   lps::PCISPHSolver  LiPhEn::PCISPHSolver  LiquidPhysics/include/PCISPHSolver.h:7-21
 *   update(dt)         PhysicSolver::update  LiquidPhysics/src/PhysicSolver.cpp:12-21
 *
 * Same method names, parameter meanings and constructor defaults
 * (SPHSolver.cpp:11-23, PCISPHSolver.cpp:6-7).  Differences, all following
 * from the device-resident state:
 *  - particles are values (position, velocity), not SPHParticle pointers that
 *    the solver writes back each step; read them with getPositions() /
 *    getVelocities() (the solver's sorted order, as the reference's readback,
 *    SPHSolver.cpp:412-422) or write them into a device buffer with
 *    copyPositionsDevice();
 *  - particles added between steps are appended on the device (the reference
 *    re-uploads all of them, SPHSolver.cpp:791); the first upload gives every
 *    particle a zero half velocity and the first-step flag;
 *  - a failed library call throws lps::Error carrying the lps_status (the
 *    reference has no error channel);
 *  - a plane collider and the PCISPH loop additions of lps_params
 *    (corrections that always run, stall rule, warm start) are exposed too.
 * Header-only; link with liblps.so.
 */
#ifndef LPS_SOLVER_HPP
#define LPS_SOLVER_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lps.h"

namespace lps {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& where)
        : std::runtime_error(where + ": liblps status " + std::to_string(status)), m_status(status) {}
    int status() const { return m_status; }

private:
    int m_status;
};

inline void check(int status, const char* where) {
    if (status != LPS_OK) throw Error(status, where);
}

struct Vector3D {
    float x = 0.f, y = 0.f, z = 0.f;
};

// StaticCollisionBox / StaticCollisionSphere (LiquidPhysics/include/Collision/),
// and a plane (fluid on the side n.x >= offset).
struct CollisionBox {
    Vector3D center, halfDimensions;
    bool isBoundary = true;
};
struct CollisionSphere {
    Vector3D center;
    float radius = 0.f;
    bool isBoundary = false;
};
struct CollisionPlane {
    Vector3D normal;
    float offset = 0.f;
};

class SPHSolver {
public:
    SPHSolver() : SPHSolver(LPS_METHOD_SPH) {}
    virtual ~SPHSolver() {
        if (m_ctx) lps_destroy(m_ctx);
    }
    SPHSolver(const SPHSolver&) = delete;
    SPHSolver& operator=(const SPHSolver&) = delete;

    // One step (PhysicSolver::update): uploads what changed since the last
    // step, then runs the whole step on the device.
    void update(float deltaTime) {
        sync();
        check(lps_step(m_ctx, deltaTime, 1, &m_stats), "lps_step");
    }

    void addParticle(const Vector3D& position, const Vector3D& velocity = Vector3D{}) {
        auto& x = m_uploaded ? m_addX : m_x;
        auto& v = m_uploaded ? m_addV : m_v;
        x.insert(x.end(), {position.x, position.y, position.z});
        v.insert(v.end(), {velocity.x, velocity.y, velocity.z});
    }
    void removeParticles() {
        m_x.clear();
        m_v.clear();
        m_addX.clear();
        m_addV.clear();
        m_uploaded = false;
        m_count = 0;
    }

    void addStaticCollisionObject(const CollisionBox& b) {
        m_boxes.push_back(lps_box{{b.center.x, b.center.y, b.center.z},
                                  {b.halfDimensions.x, b.halfDimensions.y, b.halfDimensions.z},
                                  b.isBoundary ? 1 : 0});
        m_collidersChanged = true;
    }
    void addStaticCollisionObject(const CollisionSphere& s) {
        m_spheres.push_back(lps_sphere{{s.center.x, s.center.y, s.center.z}, s.radius, s.isBoundary ? 1 : 0});
        m_collidersChanged = true;
    }
    void addStaticCollisionObject(const CollisionPlane& p) {
        m_planes.push_back(lps_plane{{p.normal.x, p.normal.y, p.normal.z}, p.offset});
        m_collidersChanged = true;
    }
    void removeStaticCollisionObjects() {
        m_boxes.clear();
        m_spheres.clear();
        m_planes.clear();
        m_collidersChanged = true;
    }
    // The reference's setHasCollisionObjectDataChanged: colliders are re-sent
    // before the next step (they are values here; to move one, remove the
    // colliders and add them again, which marks them changed as well).
    void setHasCollisionObjectDataChanged(bool changed) { m_collidersChanged = m_collidersChanged || changed; }

    // SPHSolver::cleanUp (SPHSolver.cpp:180-184): no particles, no colliders.
    void cleanUp() {
        removeParticles();
        removeStaticCollisionObjects();
    }

    int getParticleCount() const {
        return (int)(m_uploaded ? m_count + m_addX.size() / 3 : m_x.size() / 3);
    }
    // Positions / velocities (xyz interleaved) in the solver's order: after a
    // step, the sorted order of that step.
    std::vector<float> getPositions() const { return readback(0); }
    std::vector<float> getVelocities() const { return readback(1); }
    // Each particle's index in insertion order (removeParticles restarts it).
    std::vector<uint32_t> getParticleIds() const {
        std::vector<uint32_t> id;
        if (!m_uploaded) {
            id.resize(m_x.size() / 3);
            for (size_t i = 0; i < id.size(); i++) id[i] = (uint32_t)i;
            return id;
        }
        const_cast<SPHSolver*>(this)->sync();
        id.resize(m_count);
        check(lps_get_particles(m_ctx, nullptr, nullptr, nullptr, nullptr, id.data(), m_count), "lps_get_particles");
        return id;
    }
    // Positions straight into device memory (e.g. a mapped vertex buffer).
    void copyPositionsDevice(float* dst, uint32_t stride, uint64_t capacityBytes) {
        sync();
        check(lps_copy_positions_device(m_ctx, dst, stride, m_count, capacityBytes), "lps_copy_positions_device");
    }
    const lps_step_stats& getLastStepStats() const { return m_stats; }

    Vector3D getGravity() const { return Vector3D{m_params.gravity[0], m_params.gravity[1], m_params.gravity[2]}; }
    float getParticleRadius() const { return m_params.particle_radius; }
    float getParticleMass() const { return derived(2); }
    float getKernelRadius() const { return derived(0); }
    float getKernelRadiusFactor() const { return m_params.kernel_radius_factor; }
    float getRestDensity() const { return m_params.rest_density; }
    float getViscosityCoefficient() const { return m_params.viscosity; }
    float getPressureStiffnessCoefficient() const { return m_params.stiffness; }
    float getNegativePressureFactor() const { return m_params.negative_pressure_factor; }
    float getSurfaceTensionCoefficient() const { return m_params.surface_tension; }
    float getSurfaceTensionThreshold() const { return m_params.surface_tension_threshold; }
    float getResitutionCoefficient() const { return m_params.restitution; }  // reference spelling
    float getFrictionCoefficient() const { return m_params.friction; }

    void setGravity(const Vector3D& g) {
        set(m_params.gravity[0], g.x);
        set(m_params.gravity[1], g.y);
        set(m_params.gravity[2], g.z);
    }
    void setParticleRadius(float v) { set(m_params.particle_radius, v); }
    void setKernelRadiusFactor(float v) { set(m_params.kernel_radius_factor, v); }
    void setRestDensity(float v) { set(m_params.rest_density, v); }
    void setViscosityCoefficient(float v) { set(m_params.viscosity, v); }
    void setPressureStiffnessCoefficient(float v) { set(m_params.stiffness, v); }
    void setNegativePressureFactor(float v) { set(m_params.negative_pressure_factor, v); }
    void setSurfaceTensionCoefficient(float v) { set(m_params.surface_tension, v); }
    void setSurfaceTensionThreshold(float v) { set(m_params.surface_tension_threshold, v); }
    void setRestitutionCoefficient(float v) { set(m_params.restitution, v); }
    void setFrictionCoefficient(float v) { set(m_params.friction, v); }

    // All parameters as the C ABI takes them (for callers that set fields
    // the reference does not have).
    const lps_params& params() const { return m_params; }
    void setParams(const lps_params& p) {
        m_params = p;
        m_params.abi_version = LPS_ABI_VERSION;
        m_paramsChanged = true;
    }

protected:
    explicit SPHSolver(int method) {
        m_params.abi_version = LPS_ABI_VERSION;
        m_params.method = method;
        m_params.gravity[0] = 0.f;
        m_params.gravity[1] = -9.81f;
        m_params.gravity[2] = 0.f;
        m_params.particle_radius = 0.017f;
        m_params.kernel_radius_factor = 2.688f;
        m_params.rest_density = 998.29f;
        m_params.stiffness = 3.f;
        m_params.negative_pressure_factor = 0.f;
        m_params.viscosity = 3.5f;
        m_params.surface_tension = 0.0728f;
        m_params.surface_tension_threshold = 7.065f;
        m_params.restitution = 0.5f;
        m_params.friction = 1.f;
        m_params.mass_override = 0.f;
        m_params.min_iterations = 4;
        m_params.max_iterations = 4;
        m_params.max_density_error_ratio = 0.1f;
        m_params.warm_start = LPS_WARM_NONE;
        m_params.stall_iterations = 0;
    }

    void set(float& field, float v) {
        if (field != v) {
            field = v;
            m_paramsChanged = true;
        }
    }
    void seti(int32_t& field, int32_t v) {
        if (field != v) {
            field = v;
            m_paramsChanged = true;
        }
    }

    lps_params m_params{};

private:
    // Creates the context on first use and sends what changed: parameters,
    // colliders, all particles (after removeParticles or before the first
    // step) or the ones added since the last step.
    void sync() {
        if (!m_ctx) {
            check(lps_create(&m_params, &m_ctx), "lps_create");
            m_paramsChanged = false;
        }
        if (m_paramsChanged) {
            check(lps_set_params(m_ctx, &m_params), "lps_set_params");
            m_paramsChanged = false;
        }
        if (m_collidersChanged) {
            check(lps_set_colliders(m_ctx, m_boxes.data(), (uint32_t)m_boxes.size(), m_spheres.data(),
                                    (uint32_t)m_spheres.size(), m_planes.data(), (uint32_t)m_planes.size()),
                  "lps_set_colliders");
            m_collidersChanged = false;
        }
        if (!m_uploaded) {
            const uint32_t n = (uint32_t)(m_x.size() / 3);
            if (n == 0) throw Error(LPS_ERR_STATE, "update: no particles");
            check(lps_set_particles(m_ctx, m_x.data(), m_v.data(), nullptr, nullptr, n), "lps_set_particles");
            m_count = n;
            m_uploaded = true;
            m_x.clear();
            m_v.clear();
        }
        if (!m_addX.empty()) {
            const uint32_t n = (uint32_t)(m_addX.size() / 3);
            check(lps_add_particles(m_ctx, m_addX.data(), m_addV.data(), nullptr, nullptr, n), "lps_add_particles");
            m_count += n;
            m_addX.clear();
            m_addV.clear();
        }
    }

    std::vector<float> readback(int which) const {
        if (!m_uploaded) return which == 0 ? m_x : m_v;
        const_cast<SPHSolver*>(this)->sync();
        std::vector<float> out(3 * (size_t)m_count);
        check(lps_get_particles(m_ctx, which == 0 ? out.data() : nullptr, which == 1 ? out.data() : nullptr, nullptr,
                                nullptr, nullptr, m_count),
              "lps_get_particles");
        return out;
    }

    float derived(int which) const {
        const_cast<SPHSolver*>(this)->ensureContext();
        float h = 0.f, div = 0.f, mass = 0.f;
        check(lps_get_derived(m_ctx, &h, &div, &mass, nullptr), "lps_get_derived");
        return which == 0 ? h : which == 1 ? div : mass;
    }
    void ensureContext() {
        if (!m_ctx) {
            check(lps_create(&m_params, &m_ctx), "lps_create");
            m_paramsChanged = false;
            m_collidersChanged = true;
        } else if (m_paramsChanged) {
            check(lps_set_params(m_ctx, &m_params), "lps_set_params");
            m_paramsChanged = false;
        }
    }

    lps_ctx* m_ctx = nullptr;
    lps_step_stats m_stats{};
    std::vector<float> m_x, m_v;        // particles not yet on the device (all of them until the first step)
    std::vector<float> m_addX, m_addV;  // particles added after the upload
    std::vector<lps_box> m_boxes;
    std::vector<lps_sphere> m_spheres;
    std::vector<lps_plane> m_planes;
    uint32_t m_count = 0;
    bool m_uploaded = false;
    bool m_paramsChanged = false;
    bool m_collidersChanged = true;
};

// The reference's PCISPH loop runs at most m_minIterations corrections and
// breaks after the first one whose maximum density error ratio is below the
// tolerance (PCISPHSolver.cpp:64,158-159, 180 for the OpenCL path): in
// lps_params terms min_iterations = 1 and max_iterations = m_minIterations.
// getMinIterations / setMinIterations keep that meaning; the library's own
// minimum (corrections always run, as the benchmark's min 3) and the stall
// rule are separate additions.
class PCISPHSolver : public SPHSolver {
public:
    PCISPHSolver() : SPHSolver(LPS_METHOD_PCISPH) {
        m_params.min_iterations = 1;
        m_params.max_iterations = 4;  // m_minIterations = 4 (PCISPHSolver.cpp:6)
        m_params.max_density_error_ratio = 0.1f;
    }

    int getMinIterations() const { return m_params.max_iterations; }
    float getMaxDensityErrorRatio() const { return m_params.max_density_error_ratio; }
    void setMinIterations(int minIterations) {
        seti(m_params.max_iterations, minIterations);
        if (m_params.min_iterations > minIterations) seti(m_params.min_iterations, minIterations);
    }
    void setMaxDensityErrorRatio(float maxDensityErrorRatio) { set(m_params.max_density_error_ratio, maxDensityErrorRatio); }

    // Additions: corrections that always run, the stall rule (stop when the
    // error has not improved on its best for this many corrections; 0 off)
    // and the learned warm start.
    int getMinCorrections() const { return m_params.min_iterations; }
    void setMinCorrections(int n) { seti(m_params.min_iterations, n); }
    int getStallIterations() const { return m_params.stall_iterations; }
    void setStallIterations(int stallIterations) { seti(m_params.stall_iterations, stallIterations); }
    void setWarmStart(lps_warm_start mode) { seti(m_params.warm_start, (int32_t)mode); }
};

}  // namespace lps

#endif  // LPS_SOLVER_HPP
