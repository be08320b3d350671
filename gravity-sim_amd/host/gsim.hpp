// C++ surface of libgsim with the reference's class names and call shapes:
//   gsim::Exception          src/Debug/Exception.hpp:10-72 (file, line, message)
//   gsim::ParticleSystem     src/Particles/ParticleSystem.hpp:58-181
//   gsim::DirectSimulation   src/Simulation/Direct/DirectSimulation.hpp:15-63
//   gsim::BarnesHutSimulation src/Simulation/BarnesHut/BarnesHutSimulation.hpp:15-52
// Every C-ABI failure is rethrown as gsim::Exception carrying gsim_last_error().
// The device state lives in the simulation object (a gsim_sim handle); the
// particle system reads it back through the simulation attached to it.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <ctime>
#include <exception>
#include <vector>

#include "gsim.h"

namespace gsim {

class Exception : public std::exception {
   public:
    Exception(const char *file, uint32_t line, const char *format, ...) : file_(file), line_(line) {
        va_list args;
        va_start(args, format);
        std::vsnprintf(message_, sizeof(message_), format, args);
        va_end(args);
    }
    const char *GetFile() const { return file_; }
    uint32_t GetLine() const { return line_; }
    const char *GetMessage() const { return message_; }
    const char *what() const noexcept override { return message_; }

   private:
    const char *file_;
    uint32_t line_;
    char message_[512];
};

#define GSIM_THROW_EXCEPTION(format, ...) \
    throw gsim::Exception(__FILE__, __LINE__, format __VA_OPT__(, ) __VA_ARGS__)
#define GSIM_CHECK(call)                                                          \
    do {                                                                          \
        gsim_status st_ = (call);                                                 \
        if (st_ != GSIM_OK) GSIM_THROW_EXCEPTION("%s", gsim_last_error());        \
    } while (0)

class ParticleSystem {
   public:
    enum GenerateType {
        GENERATE_TYPE_RANDOM,
        GENERATE_TYPE_GALAXY,
        GENERATE_TYPE_GALAXY_COLLISION,
        GENERATE_TYPE_SYMMETRICAL_GALAXY_COLLISION,
        GENERATE_TYPE_COUNT
    };
    enum SimulationAlgorithm {
        SIMULATION_ALGORITHM_DIRECT_SUM,
        SIMULATION_ALGORITHM_BARNES_HUT,
        SIMULATION_ALGORITHM_COUNT
    };

    // ParticleSystem.cpp:438-498: particles from a text file.
    ParticleSystem(const char *filePath, float gravitationalConst, float simulationTime,
                   float simulationSpeed, float softeningLen, float accuracyParameter,
                   SimulationAlgorithm algorithm)
        : gravitationalConst_(gravitationalConst), simulationTime_(simulationTime),
          simulationSpeed_(simulationSpeed), softeningLen_(softeningLen),
          accuracyParameter_(accuracyParameter), algorithm_(algorithm) {
        CheckAlgorithm();
        size_t n = 0;
        GSIM_CHECK(gsim_count_particles(filePath, &n));
        Resize(n);
        GSIM_CHECK(gsim_load_particles(filePath, pos_.data(), vel_.data(), mass_.data(), n, &n));
    }

    // ParticleSystem.cpp:499-559: generated particles.  The reference seeds
    // with time(nullptr); the seed is explicit here (default: the time).
    ParticleSystem(size_t particleCount, GenerateType generateType, float generateSize,
                   float minMass, float maxMass, float gravitationalConst, float simulationTime,
                   float simulationSpeed, float softeningLen, float accuracyParameter,
                   SimulationAlgorithm algorithm, uint64_t seed = (uint64_t)std::time(nullptr))
        : gravitationalConst_(gravitationalConst), simulationTime_(simulationTime),
          simulationSpeed_(simulationSpeed), softeningLen_(softeningLen),
          accuracyParameter_(accuracyParameter), algorithm_(algorithm) {
        CheckAlgorithm();
        if (generateType == GENERATE_TYPE_SYMMETRICAL_GALAXY_COLLISION) particleCount &= ~(size_t)1;
        if (!particleCount) GSIM_THROW_EXCEPTION("The simulation must contain at least one particle!");
        Resize(particleCount);
        GSIM_CHECK(gsim_generate((int32_t)generateType, particleCount, generateSize, minMass,
                                 maxMass, gravitationalConst, seed, pos_.data(), vel_.data(),
                                 mass_.data()));
    }

    ParticleSystem(const ParticleSystem &) = delete;
    ParticleSystem &operator=(const ParticleSystem &) = delete;

    size_t GetParticleCount() const { return mass_.size(); }
    size_t GetAlignedParticleCount() const {
        const size_t a = gsim_required_alignment((int32_t)algorithm_);
        return (GetParticleCount() + a - 1) / a * a;
    }
    float GetGravitationalConst() const { return gravitationalConst_; }
    float GetSimulationTime() const { return simulationTime_; }
    float GetSimulationSpeed() const { return simulationSpeed_; }
    float GetSofteningLen() const { return softeningLen_; }
    float GetAccuracyParameter() const { return accuracyParameter_; }
    SimulationAlgorithm GetSimulationAlgorithm() const { return algorithm_; }

    // Host copies as generated/loaded (before any step).
    const float *InitialPositions() const { return pos_.data(); }
    const float *InitialVelocities() const { return vel_.data(); }
    const float *InitialMasses() const { return mass_.data(); }

    // ParticleSystem.cpp:561-726: current state, unpadded, original order.
    void GetParticles(float *pos_xy, float *vel_xy, float *mass) const {
        if (!sim_) GSIM_THROW_EXCEPTION("No simulation is attached to the particle system!");
        GSIM_CHECK(gsim_read(sim_, pos_xy, vel_xy, mass));
    }

    // ParticleSystem.cpp:727-750
    void SaveParticles(const char *filePath) const {
        std::vector<float> p(2 * GetParticleCount()), v(2 * GetParticleCount()),
            m(GetParticleCount());
        GetParticles(p.data(), v.data(), m.data());
        GSIM_CHECK(gsim_save_particles(filePath, p.data(), v.data(), m.data(), GetParticleCount()));
    }

    void Attach(gsim_sim *sim) { sim_ = sim; }

   private:
    void CheckAlgorithm() const {
        if (algorithm_ != SIMULATION_ALGORITHM_DIRECT_SUM &&
            algorithm_ != SIMULATION_ALGORITHM_BARNES_HUT)
            GSIM_THROW_EXCEPTION("Invalid simulation algorithm requested!");
    }
    void Resize(size_t n) {
        pos_.assign(2 * n, 0.0f);
        vel_.assign(2 * n, 0.0f);
        mass_.assign(n, 0.0f);
    }

    float gravitationalConst_, simulationTime_, simulationSpeed_, softeningLen_,
        accuracyParameter_;
    SimulationAlgorithm algorithm_;
    std::vector<float> pos_, vel_, mass_;
    gsim_sim *sim_ = nullptr;
};

// Shared body of the two simulation classes: the ctor uploads the particle
// system, RunSimulations enqueues steps without waiting, the dtor waits.
template <int32_t Algorithm>
class SimulationBase {
   public:
    static size_t GetRequiredParticleAlignment() { return gsim_required_alignment(Algorithm); }

    explicit SimulationBase(ParticleSystem *particleSystem, uint32_t groupWidth = 64,
                            int32_t device = 0)
        : particleSystem_(particleSystem) {
        gsim_params p;
        gsim_default_params(&p);
        p.gravitational_const = particleSystem->GetGravitationalConst();
        p.simulation_time = particleSystem->GetSimulationTime();
        p.simulation_speed = particleSystem->GetSimulationSpeed();
        p.softening_len = particleSystem->GetSofteningLen();
        p.accuracy_parameter = particleSystem->GetAccuracyParameter();
        p.algorithm = Algorithm;
        p.group_width = groupWidth;
        p.device = device;
        GSIM_CHECK(gsim_create(&p, particleSystem->InitialPositions(),
                               particleSystem->InitialVelocities(),
                               particleSystem->InitialMasses(),
                               particleSystem->GetParticleCount(), &sim_));
        particleSystem->Attach(sim_);
    }
    SimulationBase(const SimulationBase &) = delete;
    SimulationBase &operator=(const SimulationBase &) = delete;

    ParticleSystem *GetParticleSystem() { return particleSystem_; }
    gsim_sim *GetHandle() { return sim_; }

    void RunSimulations(uint32_t simulationCount) { GSIM_CHECK(gsim_step(sim_, simulationCount)); }
    void WaitIdle() { GSIM_CHECK(gsim_synchronize(sim_)); }

    ~SimulationBase() {
        if (sim_) {
            gsim_synchronize(sim_);
            // the particle system keeps reading through the handle until the
            // simulation goes away, as the reference's buffers outlive it
            particleSystem_->Attach(nullptr);
            gsim_destroy(sim_);
        }
    }

   protected:
    ParticleSystem *particleSystem_;
    gsim_sim *sim_ = nullptr;
};

using DirectSimulation = SimulationBase<GSIM_DIRECT_SUM>;
using BarnesHutSimulation = SimulationBase<GSIM_BARNES_HUT>;

}  // namespace gsim
