// gravity-sim: headless command-line driver with the reference's flags
// (src/Main.cpp:20-48, 167-217), defaults (:59-63), argument checks
// (:224-238), batches of 100 steps (:270-283), --benchmark output (:289-310)
// and --particles-out save (:320-321).  Rendering (the windowed loop) is not
// part of this build, so --no-graphics is required.
//
// Differences from the reference, each a reference bug not replicated:
//   --help alone prints the help text and exits (the reference falls through).
//   --no-graphics without --simulation-count is rejected (the reference tests
//   simulationCount, which is always 0, so it loops forever).
// Additions: --seed= (generation seed; the reference seeds from the clock),
// --group-width= (Barnes-Hut vote width W, 64 or 32), --device=, and
// --adaptive-box (Barnes-Hut tree over the particles' bounding square instead
// of the fixed [-500, 500) domain; beyond the reference, off by default) and
// --leaf-bits=15 (Barnes-Hut leaves on a 32768^2 grid at depth 15; beyond the
// reference, 10 by default).
#include <chrono>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <memory>

#include "gsim.hpp"

namespace {

const char *const kHelp =
    "%s, version %u.%u.%u\n"
    "Available parameters:\n"
    "\t--log-file: The file to output all logs to. If unspecified, logs will not be outputted to a file.\n"
    "\t--particles-in: The input file containing the starting variables of the particles. If not specified, the program will use the given generation parameters.\n"
    "\t--particles-out: The optional output file in which the variables of the particles will be written once the simulation is finished.\n"
    "\t--particle-count: The number of particles to generate. All generation parameters are ignored if an input file is specified.\n"
    "\t--generate-type: The variant to use for the particle system generation. One of the following options:\n"
    "\t\trandom: Randomly distributes particles within the generation confines.\n"
    "\t\tgalaxy: Simulates the approximate structure of a spiral galaxy spanning the generation confines.\n"
    "\t\tgalaxy-collision: Simulates the collision of two spiral galaxies, each spanning a third of the generation confines.\n"
    "\t\tsymmetrical-galaxy-collision: Simulates the collision of two symmetrical spiral galaxies, each spanning a third of the generation confines.\n"
    "\t--generate-size: The radius of the resulting generation's size.\n"
    "\t--min-mass: The minimum mass of the generated particles.\n"
    "\t--max-mass: The maximum mass of the generated particles.\n"
    "\t--gravitational-const: The gravitational constant used for the simulation. Defaulted to 1.\n"
    "\t--simulation-time: The time interval length, in seconds, simulated in one instance. Defaulted to 1e-3.\n"
    "\t--simulation-speed: The speed factor at which the simulation is run. Defaulted to 1.\n"
    "\t--softening-len: The softening length used to soften the extreme forces that would usually result from close interactions. Defaulted to 0.2.\n"
    "\t--accuracy-parameter: The accuracy parameter used to calibrate force approximation. Only used for Barnes-Hut simulations. Defaulted to 1.\n"
    "\t--simulation-algorithm: The simulation algorithm used to calculate the gravitational forces. One of the following options:\n"
    "\t\tdirect-sum: The direct-sum method, calculating every interaction between particles.\n"
    "\t\tbarnes-hut: The Barnes-Hut algorithm, organizing all particles in a quadtree.\n"
    "\t--simulation-count: The number of simulations to run before closing the program. No limit will be used if this parameter isn't specified.\n"
    "\t--seed: The seed of the particle generation. Defaulted to the current time.\n"
    "\t--group-width: The Barnes-Hut traversal group width, 64 or 32. Defaulted to 64.\n"
    "\t--device: The GPU to run on. Defaulted to 0.\n"
    "\t--adaptive-box: Barnes-Hut only: build the tree over the particles' bounding square\n"
    "\t\tinstead of the fixed [-500, 500) domain (no particle leaves it). Off by default.\n"
    "\t--leaf-bits: Barnes-Hut only: 10 (a 1024x1024 leaf grid) or 15 (32768x32768, depth 15).\n"
    "\t\tDefaulted to 10.\n"
    "Available options:\n"
    "\t--help: Displays the current message and exits the program.\n"
    "\t--log-detailed: Outputs non-crucial logs that might be useful for debugging or additional information.\n"
    "\t--no-graphics: Doesn't display the live positions of all particles, instead running the simulations in the background (required by this build, which has no renderer).\n"
    "\t--benchmark: Benchmarks the required runtime for all simulations. Ignored if --no-graphics isn't specified.\n";

enum Level { DEBUG = 1, INFO = 2, WARNING = 4, ERROR = 8, FATAL = 16 };

// The reference's Logger (src/Debug/Logger.cpp:10-48): a level prefix, the
// message, a newline, to the console and an optional file; FATAL aborts.
class Logger {
   public:
    Logger(const char *path, unsigned flags) : flags_(flags) {
        if (path) {
            file_ = std::fopen(path, "w");
            if (!file_)
                Log(ERROR, "Failed to open log file! All log messages will be outputted only to the console.");
        }
    }
    ~Logger() {
        if (file_) std::fclose(file_);
    }
    void Log(Level level, const char *fmt, ...) {
        if (!(flags_ & level)) return;
        va_list ap;
        va_start(ap, fmt);
        Emit(level, fmt, ap);
        va_end(ap);
    }
    void LogForced(Level level, const char *fmt, ...) {
        va_list ap;
        va_start(ap, fmt);
        Emit(level, fmt, ap);
        va_end(ap);
    }

   private:
    void Emit(Level level, const char *fmt, va_list ap) {
        char msg[2048];
        std::vsnprintf(msg, sizeof(msg), fmt, ap);
        const char *prefix = level == DEBUG ? "[DEBUG]       "
                             : level == INFO ? "[INFO]        "
                             : level == WARNING ? "[WARNING]     "
                             : level == ERROR ? "[ERROR]       "
                                              : "[FATAL ERROR] ";
        if (file_) {
            std::fprintf(file_, "%s%s\n", prefix, msg);
            std::fflush(file_);
        }
        std::fprintf(stdout, "%s%s\n", prefix, msg);
        std::fflush(stdout);
        if (level == FATAL) std::abort();
    }
    unsigned flags_;
    FILE *file_ = nullptr;
};

struct ProgramInfo {
    const char *logFile = nullptr;
    const char *particlesInFile = nullptr;
    const char *particlesOutFile = nullptr;
    uint32_t particleCount = 0;
    int generateType = gsim::ParticleSystem::GENERATE_TYPE_COUNT;
    float generateSize = 0.0f, minMass = 0.0f, maxMass = 0.0f;
    float gravitationalConst = 1.0f, simulationTime = 0.001f, simulationSpeed = 1.0f;
    float softeningLen = 0.2f, accuracyParameter = 1.0f;
    int simulationAlgorithm = gsim::ParticleSystem::SIMULATION_ALGORITHM_COUNT;
    uint64_t maxSimulationCount = UINT64_MAX;
    uint64_t seed = (uint64_t)std::time(nullptr);
    uint32_t groupWidth = 64;
    uint32_t leafBits = 10;
    int32_t device = 0;
    bool logDetailed = false, noGraphics = false, benchmark = false, adaptiveBox = false;
};

bool starts(const char *arg, const char *prefix, const char **value) {
    const size_t n = std::strlen(prefix);
    if (std::strncmp(arg, prefix, n) != 0) return false;
    *value = arg + n;
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 2 && !std::strcmp(argv[1], "--help")) {
        std::printf(kHelp, "Gravity-Sim (MI355X)", 1u, 0u, 0u);
        return 0;
    }
    ProgramInfo pi;
    const char *v = nullptr;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        if (starts(a, "--log-file=", &v)) pi.logFile = v;
        else if (starts(a, "--particles-in=", &v)) pi.particlesInFile = v;
        else if (starts(a, "--particles-out=", &v)) pi.particlesOutFile = v;
        else if (starts(a, "--particle-count=", &v)) pi.particleCount = (uint32_t)std::strtoull(v, nullptr, 10);
        else if (starts(a, "--generate-type=", &v)) {
            if (!std::strcmp(v, "random")) pi.generateType = gsim::ParticleSystem::GENERATE_TYPE_RANDOM;
            else if (!std::strcmp(v, "galaxy")) pi.generateType = gsim::ParticleSystem::GENERATE_TYPE_GALAXY;
            else if (!std::strcmp(v, "galaxy-collision")) pi.generateType = gsim::ParticleSystem::GENERATE_TYPE_GALAXY_COLLISION;
            else if (!std::strcmp(v, "symmetrical-galaxy-collision")) pi.generateType = gsim::ParticleSystem::GENERATE_TYPE_SYMMETRICAL_GALAXY_COLLISION;
        } else if (starts(a, "--generate-size=", &v)) pi.generateSize = std::strtof(v, nullptr);
        else if (starts(a, "--min-mass=", &v)) pi.minMass = std::strtof(v, nullptr);
        else if (starts(a, "--max-mass=", &v)) pi.maxMass = std::strtof(v, nullptr);
        else if (starts(a, "--gravitational-const=", &v)) pi.gravitationalConst = std::strtof(v, nullptr);
        else if (starts(a, "--simulation-time=", &v)) pi.simulationTime = std::strtof(v, nullptr);
        else if (starts(a, "--simulation-speed=", &v)) pi.simulationSpeed = std::strtof(v, nullptr);
        else if (starts(a, "--softening-len=", &v)) pi.softeningLen = std::strtof(v, nullptr);
        else if (starts(a, "--accuracy-parameter=", &v)) pi.accuracyParameter = std::strtof(v, nullptr);
        else if (starts(a, "--simulation-algorithm=", &v)) {
            if (!std::strcmp(v, "direct-sum")) pi.simulationAlgorithm = gsim::ParticleSystem::SIMULATION_ALGORITHM_DIRECT_SUM;
            else if (!std::strcmp(v, "barnes-hut")) pi.simulationAlgorithm = gsim::ParticleSystem::SIMULATION_ALGORITHM_BARNES_HUT;
        } else if (starts(a, "--simulation-count=", &v)) pi.maxSimulationCount = std::strtoull(v, nullptr, 10);
        else if (starts(a, "--seed=", &v)) pi.seed = std::strtoull(v, nullptr, 10);
        else if (starts(a, "--group-width=", &v)) pi.groupWidth = (uint32_t)std::strtoul(v, nullptr, 10);
        else if (starts(a, "--device=", &v)) pi.device = (int32_t)std::strtol(v, nullptr, 10);
        else if (starts(a, "--leaf-bits=", &v)) pi.leafBits = (uint32_t)std::strtoul(v, nullptr, 10);
        else if (!std::strcmp(a, "--log-detailed")) pi.logDetailed = true;
        else if (!std::strcmp(a, "--no-graphics")) pi.noGraphics = true;
        else if (!std::strcmp(a, "--benchmark")) pi.benchmark = true;
        else if (!std::strcmp(a, "--adaptive-box")) pi.adaptiveBox = true;
    }

    Logger logger(pi.logFile, pi.logDetailed ? (DEBUG | INFO | WARNING | ERROR | FATAL)
                                             : (WARNING | ERROR | FATAL));
    // Main.cpp:224-238
    if (!pi.particlesInFile &&
        !(pi.particleCount && pi.generateType != gsim::ParticleSystem::GENERATE_TYPE_COUNT &&
          pi.generateSize && pi.minMass && pi.maxMass))
        logger.Log(FATAL, "If a particle input file is not provided, valid generation args must be provided!");
    if (pi.particlesInFile &&
        (pi.particleCount || pi.generateType != gsim::ParticleSystem::GENERATE_TYPE_COUNT ||
         pi.generateSize || pi.minMass || pi.maxMass))
        logger.Log(WARNING, "A particle input file was provided, therefore the given generation args will be ignored.");
    if (pi.simulationAlgorithm == gsim::ParticleSystem::SIMULATION_ALGORITHM_COUNT)
        logger.Log(FATAL, "A valid simulation algorithm must be given!");
    if (!pi.noGraphics)
        logger.Log(FATAL, "This build has no renderer: run with --no-graphics and a --simulation-count.");
    if (pi.noGraphics && pi.maxSimulationCount == UINT64_MAX)
        logger.Log(FATAL, "If --no-graphics was specified, a valid simulation count must be given!");

    try {
        const auto algo = (gsim::ParticleSystem::SimulationAlgorithm)pi.simulationAlgorithm;
        std::unique_ptr<gsim::ParticleSystem> ps;
        if (pi.particlesInFile) {
            ps.reset(new gsim::ParticleSystem(pi.particlesInFile, pi.gravitationalConst,
                                              pi.simulationTime, pi.simulationSpeed,
                                              pi.softeningLen, pi.accuracyParameter, algo));
        } else {
            ps.reset(new gsim::ParticleSystem(
                pi.particleCount, (gsim::ParticleSystem::GenerateType)pi.generateType,
                pi.generateSize, pi.minMass, pi.maxMass, pi.gravitationalConst, pi.simulationTime,
                pi.simulationSpeed, pi.softeningLen, pi.accuracyParameter, algo, pi.seed));
        }
        logger.Log(INFO, "Particle system: %zu particles (%zu aligned), %s.", ps->GetParticleCount(),
                   ps->GetAlignedParticleCount(),
                   algo == gsim::ParticleSystem::SIMULATION_ALGORITHM_DIRECT_SUM ? "direct-sum"
                                                                                 : "barnes-hut");
        std::unique_ptr<gsim::DirectSimulation> ds;
        std::unique_ptr<gsim::BarnesHutSimulation> bh;
        if (algo == gsim::ParticleSystem::SIMULATION_ALGORITHM_DIRECT_SUM)
            ds.reset(new gsim::DirectSimulation(ps.get(), pi.groupWidth, pi.device));
        else
        {
            bh.reset(new gsim::BarnesHutSimulation(ps.get(), pi.groupWidth, pi.device));
            if (pi.adaptiveBox)
                GSIM_CHECK(gsim_set_tuning(bh->GetHandle(), GSIM_TUNE_BH_ADAPTIVE_BOX, 1));
            if (pi.leafBits != 10)
                GSIM_CHECK(gsim_set_tuning(bh->GetHandle(), GSIM_TUNE_BH_LEAF_BITS, pi.leafBits));
        }
        logger.Log(INFO, "Device library: %s", gsim_version());

        const auto start = std::chrono::steady_clock::now();
        // Main.cpp:270-283: batches of 100 simulations, asynchronous
        uint64_t count = 0;
        while (count != pi.maxSimulationCount) {
            uint64_t target = count + 100;
            if (target > pi.maxSimulationCount) target = pi.maxSimulationCount;
            const uint32_t n = (uint32_t)(target - count);
            if (ds) ds->RunSimulations(n);
            else bh->RunSimulations(n);
            count = target;
        }
        if (ds) ds->WaitIdle();
        else bh->WaitIdle();

        if (pi.benchmark) {  // Main.cpp:289-310
            const float runtimeSec =
                std::chrono::duration<float>(std::chrono::steady_clock::now() - start).count();
            const float runtimeAvgMs = runtimeSec * 1000 / (float)pi.maxSimulationCount;
            if (runtimeSec >= 1) logger.LogForced(INFO, "Total simulation runtime: %.3fs", runtimeSec);
            else logger.LogForced(INFO, "Total simulation runtime: %.1fms", runtimeSec * 1000);
            if (runtimeAvgMs >= 5000) logger.LogForced(INFO, "Average runtime/simulation: %.3fs", runtimeAvgMs * 0.001f);
            else if (runtimeAvgMs >= 1) logger.LogForced(INFO, "Average runtime/simulation: %.3fms", runtimeAvgMs);
            else logger.LogForced(INFO, "Average runtime/simulation: %.1fus", runtimeAvgMs * 1000);
        }
        if (pi.particlesOutFile) ps->SaveParticles(pi.particlesOutFile);
    } catch (const gsim::Exception &e) {
        logger.Log(FATAL, "%s", e.GetMessage());  // Main.cpp:423-429
    }
    return 0;
}
