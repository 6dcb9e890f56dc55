// Generated by gen_log_poly.py -- do not edit.
// P(z) ~ atanh(sqrt z)/sqrt z on [0, 0.029437252], max rel. error 1.15e-18
#pragma once
static constexpr int KF_LOG_POLY_N = 8;
static constexpr double KF_LOG_POLY[8] = {
    9.9999999999999999887e-1,
    3.3333333333333822829e-1,
    1.9999999999651169056e-1,
    1.4285714380320839941e-1,
    1.1111098528363023304e-1,
    9.0918158401146644019e-2,
    7.656264074182095483e-2,
    7.4048551803276383491e-2
};
