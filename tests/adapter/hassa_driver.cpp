// hassa_driver.cpp -- a reference-style driver (the loop body of src/main.cu:612-645)
// compiled against include/pd_mechanics_adapter.hpp instead of mechanics.cuh.
//
// Its Particle and matrix types are written here with the members the
// reference's types expose (include/SetupParticleSystem.h:11-22,
// include/matrix.h:7-26), to show that a driver using them links unchanged.
// Usage: hassa_driver <mesh.dat> <out.txt> [gpu|seq|damage]
// gpu / seq: one step through compute_velocity_GPU_host / computeVelocity;
// writes "vx vy" per particle with 17 significant digits.
// damage: computeDamageStatus (src/main_seq.cpp:473) twice at the prescribed
// displacement u_x = s (0.7 x + 0.2 y + 0.01 y^2), u_y = s (0.4 y - 0.1 x),
// s = 4e-4, with the reference's initial state (bondDamage = 0,
// bondDamageThreshold = 3 / 3.5e4, :660-672); writes "status" per particle to
// out.txt and "d kappa" per bond (neighbour-list order) to out.txt.bonds.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "pd_mechanics_adapter.hpp"

using namespace std;

struct Particle {
    int globalID;
    int partitionID;
    long double volume;
    double damageStatus;
    vector<long double> initialPositions;
    vector<long double> currentPositions;
    vector<int> neighbors;
    set<int> nneighbors;
};

struct matrix {
    int ndim = 0;
    vector<vector<long double>> elements;
    matrix() {}
    matrix(int n, const vector<long double>& v) : ndim(n), elements(n, vector<long double>(n))
    {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) elements[i][j] = v[i * n + j];
    }
};

int main(int argc, char** argv)
{
    if (argc < 3) {
        cerr << "usage: hassa_driver mesh.dat out.txt [gpu|seq]\n";
        return 2;
    }
    const bool gpu_entry = argc < 4 || string(argv[3]) == "gpu";
    const bool damage_entry = argc >= 4 && string(argv[3]) == "damage";
    // mesh with two edge notches removed (readMeshFile semantics)
    ifstream in(argv[1]);
    string line;
    getline(in, line);
    vector<Particle> ps;
    int pid = 0;
    while (getline(in, line)) {
        istringstream iss(line);
        long double x, y, a;
        if (!(iss >> x >> y >> a)) continue;
        if (fabsl(x) > 17.5L && fabsl(y) < 2.0L) continue;
        Particle p;
        p.globalID = pid++;
        p.partitionID = 0;
        p.volume = a;
        p.damageStatus = 0;
        p.initialPositions = {x, y};
        p.currentPositions = {x, y};
        ps.push_back(p);
    }
    const int n = (int)ps.size();
    const long double dx = 1.0L, horizon = 3.0L;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            long double d2 = 0;
            for (int k = 0; k < 2; ++k) {
                long double r = ps[j].initialPositions[k] - ps[i].initialPositions[k];
                d2 += r * r;
            }
            if (sqrtl(d2) < 1.01 * dx * horizon) ps[i].neighbors.push_back(j);
        }
    unordered_map<int, int> id;
    for (int i = 0; i < n; ++i) id[ps[i].globalID] = i;
    vector<int> top, bottom;
    for (int k = 0; k < n; ++k) {
        if (ps[k].neighbors.size() > 17) continue;
        long double yk = ps[k].initialPositions[1];
        if (yk < -29.5L + 0.5L) bottom.push_back(k);
        else if (yk > 29.5L - 0.5L) top.push_back(k);
    }
    vector<vector<long double>> netF(n, vector<long double>(2, 0.0)), velocity = netF, acce = netF;
    vector<vector<Particle*>> nl(n);
    vector<vector<double>> bondDamage(n + 1);
    if (damage_entry) {
        vector<vector<double>> threshold(n);
        for (int i = 0; i < n; ++i) {
            bondDamage[i].assign(ps[i].neighbors.size(), 0.0);
            threshold[i].assign(ps[i].neighbors.size(), 3.0 / 3.5e4);
            const long double x = ps[i].initialPositions[0], y = ps[i].initialPositions[1];
            const long double s = 4e-4L;
            ps[i].currentPositions[0] = x + s * (0.7L * x + 0.2L * y + 0.01L * y * y);
            ps[i].currentPositions[1] = y + s * (0.4L * y - 0.1L * x);
        }
        try {
            for (int rep = 0; rep < 2; ++rep)
                computeDamageStatus(2, 10.0, 10.0, 1.0, 3.0, threshold, ps, nl, bondDamage, id);
        } catch (const exception& e) {
            cerr << e.what() << "\n";
            return 1;
        }
        ofstream out(argv[2]), ob(string(argv[2]) + ".bonds");
        out << setprecision(17);
        ob << setprecision(17);
        for (int i = 0; i < n; ++i) {
            out << ps[i].damageStatus << "\n";
            for (size_t j = 0; j < ps[i].neighbors.size(); ++j)
                ob << bondDamage[i][j] << " " << threshold[i][j] << "\n";
        }
        cout << "particles " << n << "\n";
        return 0;
    }
    const long double stepSize = 1.0e-9L, rho = 2.5e-6L, E = 3.5e4L, nv = 0.2L;
    double pf = (double)E / (1 - pow((double)nv, 2));
    matrix C(3, {1.0, (double)nv, 0.0, (double)nv, 1.0, 0.0, 0.0, 0.0, 1 - (double)nv});
    for (auto& r : C.elements)
        for (auto& e : r) e *= pf;
    // one step: applyVelocityBC(top), updatePositions, applyFixedBC(bottom), computeVelocity
    for (int i : top) {
        velocity[i] = {0.0L, 5.0L};
        acce[i] = {0.0L, 0.0L};
    }
    for (int i = 0; i < n; ++i)
        for (int m = 0; m < 2; ++m) {
            acce[i][m] = netF[i][m] / rho;
            ps[i].currentPositions[m] += velocity[i][m] * stepSize + 0.5 * acce[i][m] * powl(stepSize, 2);
        }
    for (int i : bottom) ps[i].currentPositions = ps[i].initialPositions;
    try {
        if (gpu_entry)
            compute_velocity_GPU_host(2, 10.0, 10.0, (double)horizon, (double)dx, (double)rho, C,
                                      (double)stepSize, velocity, nl, acce, netF, ps, id, bondDamage);
        else
            computeVelocity(2, 10.0, 10.0, (double)horizon, (double)dx, rho, C, stepSize, velocity, nl,
                            acce, netF, ps, id, bondDamage);
    } catch (const exception& e) {
        cerr << e.what() << "\n";
        return 1;
    }
    ofstream out(argv[2]);
    out << setprecision(17);
    for (int i = 0; i < n; ++i) out << velocity[i][0] << " " << velocity[i][1] << "\n";
    cout << "particles " << n << " top " << top.size() << " bottom " << bottom.size() << "\n";
    return 0;
}
