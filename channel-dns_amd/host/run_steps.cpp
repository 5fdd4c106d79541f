// Runs the step loop of cdns_simulation.hpp: reads key=value run parameters and a raw initial
// velocity payload (reference layout, binary64 (re, im) pairs), runs N iterations from a fresh
// start and writes the final velocity payload.
//   run_steps params.txt velocity_in.bin nsteps velocity_out.bin
#include "cdns_simulation.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

namespace {

cdns_host::RunParams read_params(const char *path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error(std::string("cannot read ") + path);
    std::map<std::string, double> kv;
    std::string line;
    while (std::getline(in, line)) {
        const auto eq = line.find('=');
        if (eq == std::string::npos) continue;
        kv[line.substr(0, eq)] = std::stod(line.substr(eq + 1));
    }
    auto get = [&](const std::string &k) {
        const auto it = kv.find(k);
        if (it == kv.end()) throw std::runtime_error("missing parameter " + k);
        return it->second;
    };
    cdns_host::RunParams p{};
    p.NX = (int)get("NX"); p.NY = (int)get("NY"); p.NZ = (int)get("NZ");
    p.NXS = (int)get("NXS"); p.NYS = (int)get("NYS"); p.NZS = (int)get("NZS");
    p.NX2 = (int)get("NX2"); p.NY2 = (int)get("NY2"); p.NZ2 = (int)get("NZ2");
    p.XL = get("XL"); p.ZL = get("ZL"); p.THETA = get("THETA"); p.Re = get("Re");
    p.G_pre = get("G_pre"); p.a_pre0 = get("a_pre0"); p.b_pre0 = get("b_pre0");
    for (int c = 0; c < 3; ++c) {
        p.a_V[c] = get("a_V" + std::to_string(c));
        p.b_V[c] = get("b_V" + std::to_string(c));
    }
    p.INDEX_GRAD_PRE = (int)get("INDEX_GRAD_PRE");
    p.DT = get("DT");
    p.V_Microblow0 = get("V_Microblow0");
    p.precision = (int)get("precision");
    return p;
}

} // namespace

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s params.txt velocity_in.bin nsteps velocity_out.bin\n", argv[0]);
        return 2;
    }
    try {
        const cdns_host::RunParams p = read_params(argv[1]);
        cdns_host::Simulation sim(p);
        std::vector<double> u(sim.vector_doubles());
        {
            std::ifstream in(argv[2], std::ios::binary);
            in.read(reinterpret_cast<char *>(u.data()), (std::streamsize)(u.size() * sizeof(double)));
            if (!in) throw std::runtime_error("short velocity payload");
        }
        sim.initialize(u);
        sim.run(std::stoi(argv[3]));
        const std::vector<double> out = sim.velocity();
        std::ofstream o(argv[4], std::ios::binary);
        o.write(reinterpret_cast<const char *>(out.data()), (std::streamsize)(out.size() * sizeof(double)));
        double dpdx = 0, dpdz = 0;
        sim.flow_gradients(dpdx, dpdz);
        std::printf("run_steps: %lld steps, t = %.6f, dpdx = %.17g, dpdz = %.17g\n", sim.steps_done(), sim.time(),
                    dpdx, dpdz);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "run_steps: %s\n", e.what());
        return 1;
    }
}
