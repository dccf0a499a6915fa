// Drives the C++ host layer (include/lps_solver.hpp) through the reference's
// solver interface -- setters, addParticle, addStaticCollisionObject,
// update(dt), readback -- on a small jittered dam break, SPH and PCISPH, and
// compares every step with the oracle's restatement of the OpenCL path
// (oracle/orc.h) on the same inputs: positions and velocities bit-identical,
// equal correction counts.  Particles are added mid-run (appended on both
// sides).  Also checks the layer's errors: stepping with no particles throws
// lps::Error(LPS_ERR_STATE).  Prints one line per step and "ok"; exit status 0
// when every comparison holds.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "lps_solver.hpp"
#include "orc.h"

namespace {

// A block of nx*ny*nz particles at spacing s from (s, s, s), jittered by up to
// +-amp with a fixed integer hash (the same inputs go to both sides).
std::vector<float> block(int nx, int ny, int nz, float s, float amp, uint32_t seed, float y0 = 0.f) {
    std::vector<float> x;
    uint32_t h = seed * 2654435761u + 12345u;
    auto rnd = [&]() {
        h ^= h << 13;
        h ^= h >> 17;
        h ^= h << 5;
        return (float)(h & 0xFFFFFFu) / (float)0x1000000 * 2.f - 1.f;
    };
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                x.push_back(s * (float)(i + 1) + amp * rnd());
                x.push_back(y0 + s * (float)(j + 1) + amp * rnd());
                x.push_back(s * (float)(k + 1) + amp * rnd());
            }
    return x;
}

bool same_bits(const std::vector<float>& a, const std::vector<float>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}

int run(bool pcisph) {
    std::unique_ptr<lps::SPHSolver> solver(pcisph ? new lps::PCISPHSolver() : new lps::SPHSolver());
    // parameters through the reference's setters (SPHSolver.h:57-69)
    solver->setParticleRadius(0.0125f);
    solver->setKernelRadiusFactor(3.2f);
    solver->setRestDensity(1000.f);
    solver->setViscosityCoefficient(0.01f);
    solver->setPressureStiffnessCoefficient(3.f);
    if (pcisph) {
        auto* p = static_cast<lps::PCISPHSolver*>(solver.get());
        p->setMinIterations(50);           // the reference's cap
        p->setMaxDensityErrorRatio(0.01f);
        p->setMinCorrections(3);
        p->setStallIterations(3);
    }
    solver->addStaticCollisionObject(lps::CollisionBox{{0.36f, 0.22f, 0.18f}, {0.36f, 0.22f, 0.18f}, true});
    solver->addStaticCollisionObject(lps::CollisionSphere{{0.2f, 0.05f, 0.1f}, 0.03f, false});
    const std::vector<float> x0 = block(8, 10, 8, 0.02f, 1e-4f, 7u);
    for (size_t i = 0; i < x0.size(); i += 3) solver->addParticle(lps::Vector3D{x0[i], x0[i + 1], x0[i + 2]});

    // the oracle gets the same parameters (orc_params has lps_params' layout)
    orc_params op;
    static_assert(sizeof(orc_params) == sizeof(lps_params), "parameter layouts");
    std::memcpy(&op, &solver->params(), sizeof(op));
    const orc_box box{{0.36f, 0.22f, 0.18f}, {0.36f, 0.22f, 0.18f}, 1};
    const orc_sphere sph{{0.2f, 0.05f, 0.1f}, 0.03f, 0};
    std::vector<float> x = x0, v(x0.size(), 0.f), vh(x0.size(), 0.f);
    std::vector<uint8_t> first(x0.size() / 3, 1);

    const float dt = 0.002f;
    int bad = 0;
    for (int step = 0; step < 6; step++) {
        if (step == 3) {  // particles added mid-run, appended on both sides
            const std::vector<float> xa = block(3, 2, 3, 0.02f, 1e-4f, 11u, 0.3f);
            for (size_t i = 0; i < xa.size(); i += 3)
                solver->addParticle(lps::Vector3D{xa[i], xa[i + 1], xa[i + 2]}, lps::Vector3D{0.f, -0.5f, 0.f});
            for (size_t i = 0; i < xa.size(); i += 3) {
                x.insert(x.end(), {xa[i], xa[i + 1], xa[i + 2]});
                v.insert(v.end(), {0.f, -0.5f, 0.f});
                vh.insert(vh.end(), {0.f, 0.f, 0.f});
                first.push_back(1);
            }
        }
        solver->update(dt);
        orc_trace tr;
        std::memset(&tr, 0, sizeof(tr));
        const int st = orc_step_ocl_f32(&op, &box, 1, &sph, 1, nullptr, 0, dt, (int32_t)first.size(), x.data(),
                                        v.data(), vh.data(), first.data(), &tr, 4);
        const bool okx = same_bits(solver->getPositions(), x), okv = same_bits(solver->getVelocities(), v);
        const int itg = solver->getLastStepStats().iterations, ito = pcisph ? tr.iterations : 0;
        std::printf("%s step %d: n=%d corrections %d/%d positions %s velocities %s\n", pcisph ? "PCISPH" : "SPH",
                    step, solver->getParticleCount(), itg, ito, okx ? "equal" : "DIFFER", okv ? "equal" : "DIFFER");
        if (st != 0 || !okx || !okv || itg != ito || solver->getParticleCount() != (int)first.size()) bad++;
    }
    // getters through the library's derived quantities
    float h = 0.f, div = 0.f, mass = 0.f;
    orc_derived_f32(&op, &h, &div, &mass, nullptr);
    if (solver->getKernelRadius() != h || solver->getParticleMass() != mass) {
        std::printf("derived quantities differ: h %g/%g mass %g/%g\n", solver->getKernelRadius(), h,
                    solver->getParticleMass(), mass);
        bad++;
    }
    return bad;
}

}  // namespace

int main() {
    int bad = 0;
    try {
        bad += run(false);
        bad += run(true);
    } catch (const lps::Error& e) {
        std::printf("library error: %s\n", e.what());
        std::printf("FAILED\n");
        return 2;
    }
    try {
        lps::PCISPHSolver empty;
        empty.update(0.002f);
        std::printf("stepping with no particles did not throw\n");
        bad++;
    } catch (const lps::Error& e) {
        if (e.status() != LPS_ERR_STATE) {
            std::printf("unexpected status %d: %s\n", e.status(), e.what());
            bad++;
        }
    }
    std::printf("%s\n", bad ? "FAILED" : "ok");
    return bad != 0;
}
