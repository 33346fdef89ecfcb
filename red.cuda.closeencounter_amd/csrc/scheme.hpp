// scheme.hpp -- the Runge-Kutta schemes as the reference steps them (host side).
//
// Each scheme lists, per stage state r >= 1, the stage-sum form and the (stage,
// coefficient) terms that build it, the y_(n+1) combination, the error-estimate form and
// the controller constants.  Tableaux: int_rungekutta2.cu:13-15, int_rungekutta4.cu:27-32,
// int_rungekutta5.cu:24-32 (Cash-Karp), int_rungekutta8.cu:27-47 (Fehlberg 7(8)).
#pragma once

#include <vector>

#include "red_internal.hpp"

namespace red {

// One stage combination: which form, which stages feed it with which coefficient.
struct Row {
    int form = SF_AXPY;
    int nterms = 0;
    int stage[MAX_TERMS] = {0};
    double coef[MAX_TERMS] = {0.0};
    double c1 = 0.0, c2 = 0.0;
    void add(int s, double c) { stage[nterms] = s; coef[nterms] = c; nterms++; }
};

// An embedded Runge-Kutta scheme in the shape the reference steps it.
struct Scheme {
    int n_stage = 1;            // evaluations per attempt including f0
    int n_main = 1;             // stages before y_(n+1)
    std::vector<double> c;      // stage times
    std::vector<Row> rows;      // rows[r] forms stage state r (r >= 1)
    Row np1;                    // y_(n+1)
    int err_form = -1;
    Row err;
    double lambda = 1.0;
    double exponent = 1.0;
    bool clear_events_each_attempt = false;   // RKF78 (int_rungekutta8.cu:730-733)
    bool accept_on_event = false;             // RK4 / RK5 (int_rungekutta4.cu:255-262)
    bool extra_on_yout = false;               // RK4: f5 = f(t + dt, y_(n+1))
};

inline Scheme make_scheme(int integrator, bool adaptive)
{
    Scheme S;
    switch (integrator) {
    case RED_EULER: {   // int_euler.cu:57-141
        S.n_stage = 1; S.n_main = 1; S.c = {0.0};
        S.np1.form = SF_AXPY; S.np1.add(0, 1.0);
        break;
    }
    case RED_RK2: {     // int_rungekutta2.cu:13-152
        const double a[2] = {0.0, 1.0 / 2.0}, b[2] = {0.0, 1.0};
        S.n_stage = 2; S.n_main = 2; S.c = {0.0, 1.0 / 2.0};
        S.rows.resize(2);
        S.rows[1].form = SF_AXPY; S.rows[1].add(0, a[1]);
        S.np1.form = SF_AXPY; S.np1.add(1, b[1]);
        break;
    }
    case RED_RK4: {     // int_rungekutta4.cu:27-273
        const double a[4] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
        const double b[2] = {1.0 / 6.0, 1.0 / 3.0};
        S.c = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0, 1.0};
        S.n_main = 4; S.n_stage = adaptive ? 5 : 4;
        S.rows.resize(4);
        for (int r = 1; r < 4; r++) { S.rows[r].form = SF_AXPY; S.rows[r].add(r - 1, a[r]); }
        S.np1.form = SF_RK4NP1; for (int k = 0; k < 4; k++) S.np1.add(k, 0.0);
        S.np1.c1 = b[0]; S.np1.c2 = b[1];
        S.err_form = EF_RK4; S.err.add(3, 0.0); S.err.add(4, 0.0);
        S.lambda = 1.0 / 10.0; S.exponent = 1.0 / 4.0;
        S.accept_on_event = true; S.extra_on_yout = true;
        break;
    }
    case RED_RK5: {     // Cash-Karp, int_rungekutta5.cu:24-32 (tableau), restated as a working scheme
        const double a[16] = {0.0,
                              1.0 / 5.0,
                              3.0 / 40.0, 9.0 / 40.0,
                              3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0,
                              -11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0,
                              1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0};
        const double bh[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
        const double b[6] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 1.0 / 4.0};
        const int idx[6] = {0, 1, 2, 4, 7, 11};
        S.c = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
        S.n_stage = 6; S.n_main = 6;
        S.rows.resize(6);
        for (int r = 1; r < 6; r++) {
            S.rows[r].form = SF_ACCUM;
            for (int s = 0; s < r; s++) if (a[idx[r] + s] != 0.0) S.rows[r].add(s, a[idx[r] + s]);
        }
        S.np1.form = SF_ACCUM;
        for (int s = 0; s < 6; s++) if (bh[s] != 0.0) S.np1.add(s, bh[s]);
        S.err_form = EF_CK;
        for (int s = 0; s < 6; s++)
            if (bh[s] - b[s] != 0.0) S.err.add(s, bh[s] - b[s]);
        S.lambda = 1.0; S.exponent = 1.0 / 5.0;
        S.accept_on_event = true;
        break;
    }
    case RED_RKF78: {   // int_rungekutta8.cu:27-47, 476-797
        const double a[79] = {
            0.0,
            2.0 / 27.0,
            1.0 / 36.0, 1.0 / 12.0,
            1.0 / 24.0, 0.0, 1.0 / 8.0,
            5.0 / 12.0, 0.0, -25.0 / 16.0, 25.0 / 16.0,
            1.0 / 20.0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 5.0,
            -25.0 / 108.0, 0.0, 0.0, 125.0 / 108.0, -65.0 / 27.0, 125.0 / 54.0,
            31.0 / 300.0, 0.0, 0.0, 0.0, 61.0 / 225.0, -2.0 / 9.0, 13.0 / 900.0,
            2.0, 0.0, 0.0, -53.0 / 6.0, 704.0 / 45.0, -107.0 / 9.0, 67.0 / 90.0, 3.0,
            -91.0 / 108.0, 0.0, 0.0, 23.0 / 108.0, -976.0 / 135.0, 311.0 / 54.0, -19.0 / 60.0, 17.0 / 6.0, -1.0 / 12.0,
            2383.0 / 4100.0, 0.0, 0.0, -341.0 / 164.0, 4496.0 / 1025.0, -301.0 / 82.0, 2133.0 / 4100.0, 45.0 / 82.0, 45.0 / 164.0, 18.0 / 41.0,
            3.0 / 205.0, 0.0, 0.0, 0.0, 0.0, -6.0 / 41.0, -3.0 / 205.0, -3.0 / 41.0, 3.0 / 41.0, 6.0 / 41.0, 0.0,
            -1777.0 / 4100.0, 0.0, 0.0, -341.0 / 164.0, 4496.0 / 1025.0, -289.0 / 82.0, 2193.0 / 4100.0, 51.0 / 82.0, 33.0 / 164.0, 12.0 / 41.0, 0.0, 1.0};
        const double b[11] = {41.0 / 840.0, 0.0, 0.0, 0.0, 0.0, 34.0 / 105.0, 9.0 / 35.0, 9.0 / 35.0, 9.0 / 280.0, 9.0 / 280.0, 41.0 / 840.0};
        const int idx[13] = {0, 1, 2, 4, 7, 11, 16, 22, 29, 37, 46, 56, 67};
        S.c = {0.0, 2.0 / 27.0, 1.0 / 9.0, 1.0 / 6.0, 5.0 / 12.0, 1.0 / 2.0, 5.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0, 1.0, 0.0, 1.0};
        S.n_main = 11; S.n_stage = adaptive ? 13 : 11;
        S.rows.resize(13);
        for (int r = 1; r < 13; r++) {
            S.rows[r].form = SF_DTSUM;
            for (int s = 0; s < r; s++) if (a[idx[r] + s] != 0.0) S.rows[r].add(s, a[idx[r] + s]);
        }
        S.np1.form = SF_RKF78NP1;
        S.np1.add(0, b[0]); S.np1.add(5, b[5]); S.np1.add(6, b[6]); S.np1.add(7, b[7]);
        S.np1.add(8, b[8]); S.np1.add(9, b[9]); S.np1.add(10, b[10]);
        S.err_form = EF_RKF78;
        S.err.add(0, 0.0); S.err.add(10, 0.0); S.err.add(11, 0.0); S.err.add(12, 0.0);
        S.lambda = 41.0 / 840.0; S.exponent = 1.0 / (7 + 1);
        S.clear_events_each_attempt = true;
        break;
    }
    }
    return S;
}

}  // namespace red
