// oracle/ref_driver.cpp -- C entry points over the reference's own helper
// classes (TEST INFRASTRUCTURE, built into oracle/_ref only).
//
// The reference's SPHSolver.cpp / PCISPHSolver.cpp / OpenCLInterface.cpp need
// <CL/cl.hpp>, which this image does not ship, so those translation units are
// not built.  Everything they call on the sequential (NONE) path compiles from
// the reference sources unchanged: Vector3D, SPHKernel and its three kernels,
// SPHParticle/PCISPHParticle, SPHParticleEmitter, StaticCollisionBox/Sphere and
// SPHSpatialGrid.  This driver links those objects and restates only the short
// solver loops of the NONE branches (SPHSolver.cpp:188-464, 480-507;
// PCISPHSolver.cpp:55-170, 286-328) on top of them, so the oracle's FP32
// restatement can be compared with the reference's own arithmetic.
#include <cstdint>
#include <cstring>
#include <vector>
#include <cmath>
#include <algorithm>

#include "Collision/StaticCollisionBox.h"
#include "Collision/StaticCollisionSphere.h"
#include "Kernels/DefaultKernel.h"
#include "Kernels/PressureKernel.h"
#include "Kernels/ViscosityKernel.h"
#include "Particles/PCISPHParticle.h"
#include "Particles/SPHParticleEmitter.h"
#include "SPHSpatialGrid.h"

using namespace LiPhEn;

namespace {

struct RefParams {
    int32_t method;
    float gravity[3];
    float particle_radius;
    float kernel_radius_factor;
    float rest_density;
    float stiffness;
    float negative_pressure_factor;
    float viscosity;
    float surface_tension;
    float surface_tension_threshold;
    float restitution;
    float friction;
    int32_t min_iterations;
    int32_t max_iterations;
    float max_density_error_ratio;
};

struct Solver {
    DefaultKernel dk;
    PressureKernel pk;
    ViscosityKernel vk;
    float h, mass;
};

// SPHSolver::recalcKernelRadius + recalcParticleMass (SPHSolver.cpp:480-523)
void setup(Solver& s, float rp, float factor, float rho0) {
    s.h = factor * rp;
    s.dk.setRadius(s.h);
    s.pk.setRadius(s.h);
    s.vk.setRadius(s.h);
    std::vector<Vector3D> pts;
    float step = 1.6f * rp;
    for (int i = -3; i <= 3; i++)
        for (int j = -3; j <= 3; j++)
            for (int k = -3; k <= 3; k++) pts.push_back(Vector3D(i * step, j * step, k * step));
    float sum = 0.f;
    for (Vector3D q : pts) sum += s.dk.getKernelWeight(q.magnitude());
    s.mass = rho0 / sum;
}

// PCISPHSolver::calcDelta (PCISPHSolver.cpp:286-328)
float calc_delta(Solver& s, float rp, float rho0, float dt) {
    std::vector<Vector3D> pts;
    float step = 1.6f * rp;
    for (int i = -3; i <= 3; i++)
        for (int j = -3; j <= 3; j++)
            for (int k = -3; k <= 3; k++) pts.push_back(Vector3D(i * step, j * step, k * step));
    Vector3D d1(0.f, 0.f, 0.f);
    float d2 = 0.f;
    for (Vector3D q : pts) {
        float d = q.magnitude();
        if (d > 0.f) {
            Vector3D dir = q / d;
            Vector3D g = dir * s.pk.getFirstDerivativeWeight(d);
            d1 += g;
            d2 += g * g;
        }
    }
    float denom = -(d1 * d1) - d2;
    float beta = dt * s.mass / rho0;
    beta = 2.f * beta * beta;
    float delta = 0.f;
    if (fabs(denom) > 0.f) delta = -1.f / (beta * denom);
    return delta;
}

ParticleCollisionData collide(const std::vector<StaticCollisionObject*>& objs, ParticleCollisionData pd,
                              float r, float e, float mu) {
    pd.radius = r;
    pd.restitutionCoefficient = e;
    pd.frictionCoefficient = mu;
    for (StaticCollisionObject* o : objs) pd = o->handleCollisionWithParticle(pd);
    return pd;
}

std::vector<StaticCollisionObject*> make_colliders(const float* boxes, int32_t nbox, const float* spheres,
                                                   int32_t nsphere, const int32_t* kinds, const int32_t* index,
                                                   int32_t ncoll) {
    // box record: cx cy cz hx hy hz boundary ; sphere record: cx cy cz r boundary
    std::vector<StaticCollisionObject*> objs;
    for (int32_t c = 0; c < ncoll; c++) {
        if (kinds[c] == 0) {
            const float* b = boxes + 7 * index[c];
            objs.push_back(new StaticCollisionBox(Vector3D(b[0], b[1], b[2]), Vector3D(b[3], b[4], b[5]),
                                                  b[6] != 0.f ? StaticCollisionObjectType::BOUNDARY
                                                              : StaticCollisionObjectType::OBSTACLE));
        } else {
            const float* sp = spheres + 5 * index[c];
            objs.push_back(new StaticCollisionSphere(Vector3D(sp[0], sp[1], sp[2]), sp[3],
                                                     sp[4] != 0.f ? StaticCollisionObjectType::BOUNDARY
                                                                  : StaticCollisionObjectType::OBSTACLE));
        }
    }
    (void)nbox; (void)nsphere;
    return objs;
}

}  // namespace

extern "C" {

int ref_tables(float h, float* out5x1001, float* div_step) {
    DefaultKernel dk;
    PressureKernel pk;
    ViscosityKernel vk;
    dk.setRadius(h);
    pk.setRadius(h);
    vk.setRadius(h);
    std::vector<float> t[5] = {dk.getKernelWeights(), dk.getFirstDerivativeWeights(),
                               dk.getSecondDerivativeWeights(), pk.getFirstDerivativeWeights(),
                               vk.getSecondDerivativeWeights()};
    for (int k = 0; k < 5; k++) {
        if (t[k].size() != 1001) return 1;
        std::memcpy(out5x1001 + 1001 * k, t[k].data(), sizeof(float) * 1001);
    }
    if (div_step) *div_step = dk.getDivisionStep();
    return 0;
}

float ref_mass(float rp, float factor, float rho0) {
    Solver s;
    setup(s, rp, factor, rho0);
    return s.mass;
}

float ref_delta(float rp, float factor, float rho0, float dt) {
    Solver s;
    setup(s, rp, factor, rho0);
    return calc_delta(s, rp, rho0, dt);
}

int ref_spawn_cube(const float* pos, const float* half, float rp, float* out, int32_t cap) {
    std::vector<Vector3D> v = SPHParticleEmitter::spawnCube(Vector3D(pos[0], pos[1], pos[2]),
                                                            Vector3D(half[0], half[1], half[2]), rp);
    int32_t n = (int32_t)v.size();
    for (int32_t i = 0; i < n && i < cap; i++) {
        out[3 * i] = v[i].getX();
        out[3 * i + 1] = v[i].getY();
        out[3 * i + 2] = v[i].getZ();
    }
    return n;
}

// SPHParticleEmitter::spawnSphere / spawnCircleZ (SPHParticleEmitter.cpp:45-75, 137-162).
static int copy_out(const std::vector<Vector3D>& v, float* out, int32_t cap) {
    int32_t n = (int32_t)v.size();
    for (int32_t i = 0; i < n && i < cap; i++) {
        out[3 * i] = v[i].getX();
        out[3 * i + 1] = v[i].getY();
        out[3 * i + 2] = v[i].getZ();
    }
    return n;
}

int ref_spawn_sphere(const float* pos, float radius, float rp, float* out, int32_t cap) {
    return copy_out(SPHParticleEmitter::spawnSphere(Vector3D(pos[0], pos[1], pos[2]), radius, rp), out, cap);
}

int ref_spawn_circle_z(const float* pos, float radius, float rp, float* out, int32_t cap) {
    return copy_out(SPHParticleEmitter::spawnCircleZ(Vector3D(pos[0], pos[1], pos[2]), radius, rp), out, cap);
}

// Colliders applied in insertion order with the reference's collision classes.
int ref_collide(const float* boxes, int32_t nbox, const float* spheres, int32_t nsphere, const int32_t* kinds,
                const int32_t* index, int32_t ncoll, float r, float e, float mu, int32_t n, float* x, float* vel) {
    auto objs = make_colliders(boxes, nbox, spheres, nsphere, kinds, index, ncoll);
    for (int32_t i = 0; i < n; i++) {
        ParticleCollisionData pd;
        pd.position = Vector3D(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        pd.velocity = Vector3D(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        pd = collide(objs, pd, r, e, mu);
        x[3 * i] = pd.position.getX(); x[3 * i + 1] = pd.position.getY(); x[3 * i + 2] = pd.position.getZ();
        vel[3 * i] = pd.velocity.getX(); vel[3 * i + 1] = pd.velocity.getY(); vel[3 * i + 2] = pd.velocity.getZ();
    }
    for (auto* o : objs) delete o;
    return 0;
}

// SPHSpatialGrid neighbour list of particle i (SPHSolver.cpp:431-440).
int ref_neighbors(float h, int32_t n, const float* x, int32_t i, int32_t* out, int32_t cap) {
    std::vector<SPHParticle*> ps(n);
    for (int32_t k = 0; k < n; k++) {
        ps[k] = new SPHParticle();
        ps[k]->setPosition(Vector3D(x[3 * k], x[3 * k + 1], x[3 * k + 2]));
    }
    SPHSpatialGrid g(h);
    g.build(ps);
    std::vector<SPHParticle*> nb = g.findNeighborParticles(ps[i], h);
    int32_t c = (int32_t)nb.size();
    for (int32_t k = 0; k < c && k < cap; k++) {
        int32_t idx = -1;
        for (int32_t q = 0; q < n; q++)
            if (ps[q] == nb[k]) { idx = q; break; }
        out[k] = idx;
    }
    for (auto* p : ps) delete p;
    return c;
}

// NONE-path PhysicSolver::update (PhysicSolver.cpp:12-21) for `nsteps` steps on
// the reference's particle, kernel, grid and collider classes.  Output arrays
// (rho, pressure, force of the LAST step; may be NULL) are in particle order.
int ref_seq_run(const RefParams* P, const float* boxes, int32_t nbox, const float* spheres, int32_t nsphere,
                const int32_t* kinds, const int32_t* index, int32_t ncoll, float dt, int32_t nsteps, int32_t n,
                float* x, float* v, float* vh, uint8_t* first, float* rho_out, float* p_out, float* f_out,
                int32_t* iterations_out) {
    Solver s;
    setup(s, P->particle_radius, P->kernel_radius_factor, P->rest_density);
    const float m = s.mass, rho0 = P->rest_density;
    Vector3D gravity(P->gravity[0], P->gravity[1], P->gravity[2]);
    auto objs = make_colliders(boxes, nbox, spheres, nsphere, kinds, index, ncoll);
    std::vector<PCISPHParticle*> ps(n);
    for (int32_t i = 0; i < n; i++) {
        ps[i] = new PCISPHParticle();
        ps[i]->setPosition(Vector3D(x[3 * i], x[3 * i + 1], x[3 * i + 2]));
        ps[i]->setVelocity(Vector3D(v[3 * i], v[3 * i + 1], v[3 * i + 2]));
        ps[i]->setHalfVelocity(Vector3D(vh[3 * i], vh[3 * i + 1], vh[3 * i + 2]));
        ps[i]->setIsFirstTimeStep(first[i] != 0);
    }
    std::vector<SPHParticle*> base(ps.begin(), ps.end());
    for (int32_t step = 0; step < nsteps; step++) {
        // onBeginUpdate NONE: neighbour lists (SPHSolver.cpp:431-440)
        SPHSpatialGrid grid(s.h);
        grid.build(base);
        std::vector<std::vector<SPHParticle*>> lists;
        for (int32_t i = 0; i < n; i++) lists.push_back(grid.findNeighborParticles(ps[i], s.h));
        // calcParticleDensityPressure NONE (SPHSolver.cpp:444-464)
        for (int32_t i = 0; i < n; i++) {
            SPHParticle* pa = ps[i];
            float ws = 0.f;
            for (SPHParticle* nb : lists[i]) ws += s.dk.getKernelWeight((nb->getPosition() - pa->getPosition()).magnitude());
            pa->setDensity(m * ws);
            float pr = P->stiffness * (pa->getDensity() - rho0);
            if (pr < 0.f) pr *= P->negative_pressure_factor;
            pa->setPressure(pr);
        }
        // accumulateNonPressureForces NONE (SPHSolver.cpp:203-248)
        for (int32_t i = 0; i < n; i++) {
            SPHParticle* pa = ps[i];
            pa->addForce(gravity * pa->getDensity());
            Vector3D sn;
            for (SPHParticle* nb : lists[i]) {
                Vector3D dir = (pa->getPosition() - nb->getPosition());
                float d = (nb->getPosition() - pa->getPosition()).magnitude();
                sn += dir * s.dk.getFirstDerivativeWeight(d) / nb->getDensity();
            }
            sn *= m;
            float snl = sn.magnitude();
            if (snl > P->surface_tension_threshold) {
                float lc = 0.f;
                for (SPHParticle* nb : lists[i]) {
                    float d = (pa->getPosition() - nb->getPosition()).magnitude();
                    lc += s.dk.getSecondDerivativeWeight(d) / nb->getDensity();
                }
                pa->addForce((sn / snl) * (-P->surface_tension * lc * m));
            }
            Vector3D vf;
            for (SPHParticle* nb : lists[i]) {
                if (nb != pa) {
                    float d = (pa->getPosition() - nb->getPosition()).magnitude();
                    vf += ((nb->getVelocity() - pa->getVelocity()) / nb->getDensity()) * s.vk.getSecondDerivativeWeight(d);
                }
            }
            vf *= P->viscosity * m;
            pa->addForce(vf);
        }
        if (P->method == 0) {
            // accumulatePressureForces NONE (SPHSolver.cpp:270-291)
            for (int32_t i = 0; i < n; i++) {
                SPHParticle* pa = ps[i];
                Vector3D pf;
                float tf = pa->getPressure() / (pa->getDensity() * pa->getDensity());
                for (SPHParticle* nb : lists[i]) {
                    if (nb != pa) {
                        float d = (pa->getPosition() - nb->getPosition()).magnitude();
                        Vector3D dir = (pa->getPosition() - nb->getPosition()) / d;
                        pf += dir * s.pk.getFirstDerivativeWeight(d) *
                              (tf + nb->getPressure() / (nb->getDensity() * nb->getDensity()));
                    }
                }
                pf *= -(m * pa->getDensity());
                pa->addForce(pf);
            }
            if (iterations_out) iterations_out[step] = 0;
        } else {
            // PCISPHSolver::accumulatePressureForces NONE (PCISPHSolver.cpp:55-170)
            float delta = calc_delta(s, P->particle_radius, rho0, dt);
            for (int32_t i = 0; i < n; i++) {
                ps[i]->setPressure(0.f);
                ps[i]->setPredictedPressureForce(Vector3D(0.f, 0.f, 0.f));
            }
            int32_t k;
            for (k = 1; k <= P->max_iterations; k++) {
                for (int32_t i = 0; i < n; i++) {
                    PCISPHParticle* pa = ps[i];
                    Vector3D hv = pa->getHalfVelocity();
                    Vector3D acc = (pa->getAccumulatedForces() + pa->getPredictedPressureForce()) / pa->getDensity();
                    if (pa->getIsFirstTimeStep()) hv = pa->getVelocity() - acc * dt / 2.f;
                    Vector3D phv = hv + acc * dt;
                    Vector3D ppos = pa->getPosition() + phv * dt;
                    pa->setPredictedHalfVelocity(phv);
                    pa->setPredictedPosition(ppos);
                }
                for (int32_t i = 0; i < n; i++) {
                    PCISPHParticle* pa = ps[i];
                    ParticleCollisionData pd;
                    pd.position = pa->getPredictedPosition();
                    pd.velocity = pa->getPredictedHalfVelocity();
                    pd = collide(objs, pd, P->particle_radius, P->restitution, P->friction);
                    pa->setPredictedPosition(pd.position);
                    pa->setPredictedHalfVelocity(pd.velocity);
                    pa->aproximateVelocity();
                }
                for (int32_t i = 0; i < n; i++) {
                    PCISPHParticle* pa = ps[i];
                    float ws = 0.f;
                    for (SPHParticle* nb : lists[i]) {
                        PCISPHParticle* q = static_cast<PCISPHParticle*>(nb);
                        ws += s.dk.getKernelWeight((q->getPredictedPosition() - pa->getPredictedPosition()).magnitude());
                    }
                    float prd = m * ws;
                    float err = prd - rho0;
                    float pp = delta * err;
                    if (pp < 0.f) {
                        err *= P->negative_pressure_factor;
                        pp *= P->negative_pressure_factor;
                    }
                    pa->setPredictedDensity(prd);
                    pa->setDensityError(err);
                    pa->setPressure(pa->getPressure() + pp);
                }
                for (int32_t i = 0; i < n; i++) {
                    PCISPHParticle* pa = ps[i];
                    Vector3D pf;
                    float tf = pa->getPressure() / (pa->getPredictedDensity() * pa->getPredictedDensity());
                    for (SPHParticle* nb : lists[i]) {
                        PCISPHParticle* q = static_cast<PCISPHParticle*>(nb);
                        if (q != pa) {
                            float d = (q->getPredictedPosition() - pa->getPredictedPosition()).magnitude();
                            Vector3D dir = (pa->getPredictedPosition() - q->getPredictedPosition()) / d;
                            pf += dir * s.pk.getFirstDerivativeWeight(d) *
                                  (tf + nb->getPressure() / (q->getPredictedDensity() * q->getPredictedDensity()));
                        }
                    }
                    pf *= -(m * pa->getDensity());
                    if (std::isnan(pf.getX())) pf = Vector3D(0.f, 0.f, 0.f);
                    pa->setPredictedPressureForce(pf);
                }
                float mx = 0.f;
                for (PCISPHParticle* pa : ps) mx = std::max(mx, fabsf(pa->getDensityError()));
                float ratio = mx / rho0;
                if (k >= P->min_iterations && fabsf(ratio) < P->max_density_error_ratio) break;
            }
            if (k > P->max_iterations) k = P->max_iterations;
            if (iterations_out) iterations_out[step] = k;
            for (int32_t i = 0; i < n; i++) ps[i]->addForce(ps[i]->getPredictedPressureForce());
        }
        if (step == nsteps - 1) {
            for (int32_t i = 0; i < n; i++) {
                if (rho_out) rho_out[i] = ps[i]->getDensity();
                if (p_out) p_out[i] = ps[i]->getPressure();
                if (f_out) {
                    Vector3D f = ps[i]->getAccumulatedForces();
                    f_out[3 * i] = f.getX(); f_out[3 * i + 1] = f.getY(); f_out[3 * i + 2] = f.getZ();
                }
            }
        }
        // integrate + handleCollisions NONE (SPHSolver.cpp:313-355)
        for (int32_t i = 0; i < n; i++) ps[i]->integrate(dt);
        for (int32_t i = 0; i < n; i++) {
            ParticleCollisionData pd;
            pd.position = ps[i]->getPosition();
            pd.velocity = ps[i]->getHalfVelocity();
            pd = collide(objs, pd, P->particle_radius, P->restitution, P->friction);
            ps[i]->setPosition(pd.position);
            ps[i]->setHalfVelocity(pd.velocity);
            ps[i]->aproximateVelocity();
        }
    }
    for (int32_t i = 0; i < n; i++) {
        Vector3D a = ps[i]->getPosition(), b = ps[i]->getVelocity(), c = ps[i]->getHalfVelocity();
        x[3 * i] = a.getX(); x[3 * i + 1] = a.getY(); x[3 * i + 2] = a.getZ();
        v[3 * i] = b.getX(); v[3 * i + 1] = b.getY(); v[3 * i + 2] = b.getZ();
        vh[3 * i] = c.getX(); vh[3 * i + 1] = c.getY(); vh[3 * i + 2] = c.getZ();
        first[i] = ps[i]->getIsFirstTimeStep() ? 1 : 0;
        delete ps[i];
    }
    for (auto* o : objs) delete o;
    return 0;
}

}  // extern "C"
