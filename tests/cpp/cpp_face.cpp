// The C++ face (include/sucrose.hpp) driving the device the way the
// reference's C++ classes are driven: a velocity-Verlet run of 108
// Lennard-Jones atoms with the whole pair loop (calculate_all), then one
// step whose pair loop goes through calculate_forces(i, j) pair by pair over
// all pairs I < J.  Prints the energies; exit status 2 when no device.
#include <cmath>
#include <cstdio>
#include <vector>

#include "sucrose.hpp"

int main()
{
    try {
        sucrose::Device dev(0);
        const int m = 3, n = 4 * m * m * m;   // 108 atoms
        dev.set_num_atoms(n);
        if (sx_generate_fcc_lattice(dev.ctx(), m, 0.8, 0) != SX_OK) return 3;
        double box[3];
        sx_get_box(dev.ctx(), box);
        dev.set_box(box[0], box[1], box[2], true);
        dev.initialize_system(true);
        sx_set_velocity_to_temperature(dev.ctx(), 1.0, 7, 1);
        dev.init_nlist(0.3, 2.5);
        sucrose::LennardJones lj(dev, 1.0, 1.0, 2.5);
        sucrose::VerletIntegrator vv(dev, 0.004);
        dev.initialize_forces();
        lj.calculate_all();
        for (int t = 0; t < 20; ++t) {
            vv.initial_step();
            dev.update_nlist();
            dev.initialize_forces();
            lj.calculate_all();
            vv.final_step();
            dev.calculate_thermo();
        }
        // the whole loop over a list rebuilt at the current positions (the
        // reference's rebuild rule may leave pairs that came within the
        // cutoff since the last build out of the list)
        if (sx_nlist_build(dev.ctx()) != SX_OK) return 3;
        dev.initialize_forces();
        lj.calculate_all();
        double pe_loop = dev.scalar(SX_POT_ENG);
        std::vector<double> fx_loop = dev.download(SX_FX);
        // the same forces pair by pair through calculate_forces(i, j)
        dev.initialize_forces();
        for (int i = 0; i < n - 1; ++i)
            for (int j = i + 1; j < n; ++j) lj.calculate_forces(i, j);
        double pe_pairs = dev.scalar(SX_POT_ENG);
        std::vector<double> fx_pairs = dev.download(SX_FX);
        double worst = 0.0, scale = 0.0;
        for (int i = 0; i < n; ++i) {
            worst = std::fmax(worst, std::fabs(fx_pairs[i] - fx_loop[i]));
            scale = std::fmax(scale, std::fabs(fx_loop[i]));
        }
        std::printf("pe_loop %.17g pe_pairs %.17g fx_maxdiff %.3g fx_scale %.3g\n", pe_loop, pe_pairs, worst, scale);
        bool ok = std::fabs(pe_pairs - pe_loop) <= 1e-10 * std::fabs(pe_loop) && worst <= 1e-10 * scale;
        std::printf("%s\n", ok ? "CPP FACE OK" : "CPP FACE MISMATCH");
        return ok ? 0 : 1;
    } catch (const sucrose::Error &e) {
        std::printf("sucrose error %d: %s\n", e.code, e.what());
        return 2;
    }
}
