#include "params.h"

#include <cmath>
#include <fstream>
#include <locale>
#include <sstream>

#include "json.h"

namespace lfdm {

namespace {

const double kPi = 3.141592653589793238462643383279;
const double kRpm2Rad = kPi / 30;  // utils.h:14

struct Req {
    const JValue& d;
    std::string file;
    std::string& err;
    bool ok = true;

    const JValue* get(const char* key, JValue::Kind kind) {
        const JValue* v = d.find(key);
        if (!v) {
            if (ok) err = file + ": missing key \"" + key + "\"";
            ok = false;
            return nullptr;
        }
        if (v->kind != kind) {
            if (ok) err = file + ": key \"" + key + "\" has the wrong type";
            ok = false;
            return nullptr;
        }
        return v;
    }
    double num(const char* key) {
        const JValue* v = get(key, JValue::Number);
        return v ? v->num : 0.0;
    }
    bool flag(const char* key) {
        const JValue* v = get(key, JValue::Bool);
        return v ? v->b : false;
    }
    // Array of [x, y] pairs, x scaled by `sx`.
    bool pairs(const char* key, std::vector<double>& out, double sx) {
        const JValue* v = get(key, JValue::Array);
        if (!v) return false;
        out.push_back((double)v->arr.size());
        for (const JValue& e : v->arr) {
            if (e.kind != JValue::Array || e.arr.size() < 2 || e.arr[0].kind != JValue::Number ||
                e.arr[1].kind != JValue::Number) {
                if (ok) err = file + ": key \"" + key + "\" must hold [x, y] pairs";
                ok = false;
                return false;
            }
            out.push_back(e.arr[0].num * sx);
            out.push_back(e.arr[1].num);
        }
        return true;
    }
};

// tireInit (dof18.cpp:231-245): nominal and critical vertical force of the dynamic radius.
void tire_init(double pn, double rPn, double rP2n, double& fzRdynco, double& rdyncoCrit) {
    fzRdynco = (pn * (rP2n - 2.0 * rPn + 1.)) / (2. * (rP2n - rPn));
    rdyncoCrit = rPn + (rP2n - rPn) * (fzRdynco / pn - 1.);
}

// ETRTO load index -> maximum load per tire in kg, index 0..279 (GetTireMaxLoad, dof18.cpp:1465-1497).
const double kLoadIndexKg[] = {
    45, 46.5, 47.5, 48.7, 50, 51.5, 53, 54.5, 56, 58,
    60, 61.5, 63, 65, 67, 69, 71, 73, 75, 77.5,
    80.0, 82.5, 85.0, 87.5, 90.0, 92.5, 95.0, 97.5, 100.0, 103,
    106, 109, 112, 115, 118, 121, 125, 128, 132, 136,
    140, 145, 150, 155, 160, 165, 170, 175, 180, 185,
    190, 195, 200, 206, 212, 218, 224, 230, 236, 243,
    250, 257, 265, 272, 280, 290, 300, 307, 315, 325,
    335, 345, 355, 365, 375, 387, 400, 412, 425, 437,
    450, 462, 475, 487, 500, 515, 530, 545, 560, 580,
    600, 615, 630, 650, 670, 690, 710, 730, 750, 775,
    800, 825, 850, 875, 900, 925, 950, 975, 1000, 1030,
    1060, 1090, 1120, 1150, 1180, 1215, 1250, 1285, 1320, 1360,
    1400, 1450, 1500, 1550, 1600, 1650, 1700, 1750, 1800, 1850,
    1900, 1950, 2000, 2060, 2120, 2180, 2240, 2300, 2360, 2430,
    2500, 2575, 2650, 2725, 2800, 2900, 3000, 3075, 3150, 3250,
    3350, 3450, 3550, 3650, 3750, 3875, 4000, 4125, 4250, 4375,
    4500, 4625, 4750, 4875, 5000, 5150, 5300, 5450, 5600, 5850,
    6000, 6150, 6300, 6500, 6700, 6900, 7100, 7300, 7500, 7750,
    8000, 8250, 8500, 8750, 9000, 9250, 9500, 9750, 10000, 10300,
    10600, 10900, 11200, 11500, 11800, 12150, 12500, 12850, 13200, 13600,
    14000, 14500, 15000, 15550, 16000, 16500, 17000, 17500, 18000, 18500,
    19000, 19500, 20000, 20600, 21200, 21800, 22400, 23000, 23600, 24300,
    25000, 25750, 26500, 27250, 28000, 29000, 30000, 30750, 31500, 32500,
    33500, 34500, 35500, 36500, 37500, 38750, 40000, 41250, 42500, 43750,
    45000, 46250, 47500, 48750, 50000, 51500, 53000, 54500, 56000, 58000,
    60000, 61500, 63000, 65000, 67000, 69000, 71000, 73000, 75000, 77500,
    80000, 82500, 85000, 87500, 90000, 92500, 95000, 97500, 100000, 103000,
    106000, 109000, 112000, 115000, 118000, 121000, 125000, 128500, 132000, 136000,
};

double tire_max_load(unsigned li) {
    const unsigned nw = sizeof(kLoadIndexKg) / sizeof(double);
    const double g = 9.81;
    return (li < nw ? kLoadIndexKg[li] : kLoadIndexKg[nw - 1]) * g;
}

// Normalized TMeasy characteristics per tire class (GuessTruck80Par / GuessPassCar70Par,
// dof18.cpp:1500-1612): per-unit-load slopes and forces at pn and 2 pn, and slip points.
struct TireClass {
    double dfx0Pn, dfx0P2n, fxmPn, fxmP2n, fxsPn, fxsP2n, sxmPn, sxmP2n, sxsPn, sxsP2n;
    double dfy0Pn, dfy0P2n, fymPn, fymP2n, fysPn, fysP2n, symPn, symP2n, sysPn, sysP2n;
};
const TireClass kTruck80 = {17.7764, 14.5301, 0.89965, 0.77751, 0.46183, 0.42349, 0.10811, 0.12389, 0.66667, 0.66667,
                            7.4013,  6.8505,  0.75876, 0.72628, 0.68276, 0.65319, 0.33167, 0.33216, 1.0296,  1.0296};
const TireClass kPassCar70 = {18.3741, 19.4669, 1.1292, 1.0896,  0.80149, 0.76917, 0.13913, 0.13913, 0.66667, 0.66667,
                              15.9826, 12.8509, 1.0009, 0.91367, 0.8336,  0.77336, 0.14852, 0.18504, 0.96524, 1.0714};

struct NrTire {
    double jw, rr, mu = 0.8, r0, width = 0.245, rim_radius = 0.254, pn = 8562.8266, pnmax = 29969.893, kt = 411121.0;
    double rdyncoPn = 0.375, rdyncoP2n = 0.75, fzRdynco = 0, rdyncoCrit = 0;
    double c[20] = {151447.29, 236412.79, 7575.3614, 12808.276, 4657.9208, 8625.3352, 0.12, 0.15, 0.9, 0.95,
                    50931.693, 94293.847, 6615.0404, 12509.947, 6091.5092, 11443.875, 0.38786, 0.38786, 0.82534,
                    0.91309};
    double vcoulomb = 1.0, frblend_begin = 1., frblend_end = 3., bearingCapacity = 10000, li = 90, p_li = 1., p_use = 1.;
};

void guess_tire(const TireClass& k, double tireLoad, double tireWidth, double ratio, double rimDia, double p_li,
                double p_use, NrTire& t) {
    double secth = tireWidth * ratio;
    double defl_max = 0.16 * secth;
    t.pn = 0.5 * tireLoad * std::pow(p_use / p_li, 0.8);
    t.pnmax = 3.5 * t.pn;
    t.kt = tireLoad / defl_max;
    t.rim_radius = 0.5 * rimDia;
    t.width = tireWidth;
    const double pn = t.pn;
    t.c[0] = k.dfx0Pn * pn;  t.c[1] = k.dfx0P2n * 2.0 * pn;
    t.c[2] = k.fxmPn * pn;   t.c[3] = k.fxmP2n * 2.0 * pn;
    t.c[4] = k.fxsPn * pn;   t.c[5] = k.fxsP2n * 2.0 * pn;
    t.c[6] = k.sxmPn;        t.c[7] = k.sxmP2n;
    t.c[8] = k.sxsPn;        t.c[9] = k.sxsP2n;
    t.c[10] = k.dfy0Pn * pn; t.c[11] = k.dfy0P2n * 2.0 * pn;
    t.c[12] = k.fymPn * pn;  t.c[13] = k.fymP2n * 2.0 * pn;
    t.c[14] = k.fysPn * pn;  t.c[15] = k.fysP2n * 2.0 * pn;
    t.c[16] = k.symPn;       t.c[17] = k.symP2n;
    t.c[18] = k.sysPn;       t.c[19] = k.sysP2n;
}

const char* kCurveKeys[20] = {"dfx0Pn", "dfx0P2n", "fxmPn", "fxmP2n", "fxsPn", "fxsP2n", "sxmPn",
                              "sxmP2n", "sxsPn",   "sxsP2n", "dfy0Pn", "dfy0P2n", "fymPn", "fymP2n",
                              "fysPn",  "fysP2n",  "symPn", "symP2n", "sysPn",  "sysP2n"};

// Which curve keys the 11-DOF loaders double: forces and slopes (the single-track model's axle
// tire stands for the two tires of an axle, dof11.cpp:1036-1086), not the slip points.
const bool kCurveAxle[20] = {1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0};

// Low-level TMeasyNr keys (dof18.cpp:1393-1459, dof11.cpp:1208-1240 with the axle factor s2 = 2);
// mu, vcoulomb and blending keep their defaults.
void nr_low_level(Req& r, NrTire& t, double s2) {
    t.jw = r.num("jw") * s2; t.rr = r.num("rr") * s2; t.r0 = r.num("r0"); t.pn = r.num("pn") * s2;
    t.pnmax = r.num("pnmax") * s2; t.kt = r.num("kt") * s2; t.rdyncoPn = r.num("rdyncoPn");
    t.rdyncoP2n = r.num("rdyncoP2n"); t.fzRdynco = r.num("fzRdynco") * s2; t.rdyncoCrit = r.num("rdyncoCrit") * s2;
    for (int i = 0; i < 20; ++i) t.c[i] = r.num(kCurveKeys[i]) * (kCurveAxle[i] ? s2 : 1.0);
}

// Optional inflation pressures (dof18.cpp:1326-1340): both must be positive, else 1/1.
void pressures(const JValue& d, double& p_li, double& p_use) {
    const JValue* a = d.find("inflationPressureDesign");
    const JValue* b = d.find("inflationPressureUse");
    p_li = (a && a->kind == JValue::Number) ? a->num : 0.0;
    p_use = (b && b->kind == JValue::Number) ? b->num : 0.0;
    if (!(p_use > 0.0 && p_li > 0.0)) p_li = p_use = 1.0;
}

// setSuspensionParamsJSON (dof24.cpp:2192-2216): strut stiffness and damping.
bool read_suspension(const std::string& sus, std::vector<double>& flat, std::string& err) {
    if (sus.empty()) {
        err = "the 24-DOF model needs a suspension JSON file";
        return false;
    }
    JValue d;
    if (!json_load_file(sus, d, err)) return false;
    Req r{d, sus, err};
    const double ks = r.num("ks"), bs = r.num("bs");
    if (!r.ok) return false;
    flat.push_back(ks);
    flat.push_back(bs);
    return true;
}

}  // namespace

bool params_from_json(int dof, const std::string& veh, const std::string& tire, const std::string& sus, int tire_nr,
                      std::vector<double>& flat, std::string& err) {
    flat.clear();
    if (dof != 11 && dof != 18 && dof != 24) {
        err = "unsupported model: " + std::to_string(dof) + "-DOF";
        return false;
    }
    JValue d;
    if (!json_load_file(veh, d, err)) return false;
    Req r{d, veh, err};
    // setVehParamsJSON: dof11.cpp:913-1032, dof18.cpp:1085-1215, dof24.cpp:2217-2343
    static const char* const scal11[] = {"a", "b", "h", "m", "jz", "muf", "mur"};
    static const char* const scal18[] = {"a",   "b",   "h",    "m",    "jz",   "jx",   "jxz",  "cf",   "cr",
                                         "muf", "mur", "hrcf", "hrcr", "krof", "kror", "brof", "bror"};
    static const char* const scal24[] = {"a", "b", "h", "m", "jz", "jx", "jy", "cf", "cr", "muf", "mur", "hrcf", "hrcr"};
    const char* const* scal = dof == 11 ? scal11 : (dof == 18 ? scal18 : scal24);
    const int nscal = dof == 11 ? 7 : (dof == 18 ? 17 : 13);
    for (int i = 0; i < nscal; ++i) flat.push_back(r.num(scal[i]));
    bool nls = r.flag("nonLinearSteer");
    flat.push_back(r.num("maxSteer"));  // read unconditionally (dof18.cpp:1138)
    bool tc = r.flag("tcBool");
    flat.push_back(tc ? r.num("crankInertia") : 1.1);  // VehicleParam default when there is no converter
    flat.push_back(r.num("maxBrakeTorque"));
    bool thrmod = r.flag("throttleMod");
    bool drive4 = true, rear = true;
    if (d.has("4wd")) {
        drive4 = r.flag("4wd");
        rear = r.flag("rearWheels");
    }
    flat.push_back(nls ? 1 : 0);
    flat.push_back(tc ? 1 : 0);
    flat.push_back(thrmod ? 1 : 0);
    flat.push_back(drive4 ? 1 : 0);
    flat.push_back(rear ? 1 : 0);
    const JValue* g = r.get("gearRatios", JValue::Array);
    if (!r.ok) return false;
    flat.push_back((double)g->arr.size());
    for (const JValue& e : g->arr) {
        if (e.kind != JValue::Number) { err = veh + ": gearRatios must be numbers"; return false; }
        flat.push_back(e.num);
    }
    const size_t ng = g->arr.size();
    if (d.has("upshiftRPM") && d.has("downshiftRPM")) {
        double dn = r.num("downshiftRPM") * kRpm2Rad, up = r.num("upshiftRPM") * kRpm2Rad;
        flat.push_back((double)ng);
        for (size_t i = 0; i < ng; ++i) { flat.push_back(dn); flat.push_back(up); }
    } else {
        const JValue* sm = r.get("shiftMap", JValue::Array);
        if (!r.ok) return false;
        if (sm->arr.size() < ng) { err = veh + ": shiftMap needs one row per gear"; return false; }
        flat.push_back((double)ng);
        for (size_t i = 0; i < ng; ++i) {
            const JValue& e = sm->arr[i];
            if (e.kind != JValue::Array || e.arr.size() < 2) { err = veh + ": bad shiftMap row"; return false; }
            flat.push_back(e.arr[0].num * kRpm2Rad);
            flat.push_back(e.arr[1].num * kRpm2Rad);
        }
    }
    if (nls) r.pairs("steerMap", flat, 1.0); else flat.push_back(0.0);
    r.pairs("torqueMap", flat, kRpm2Rad);
    r.pairs("lossesMap", flat, kRpm2Rad);
    if (tc) {
        r.pairs("capacityFactorMap", flat, 1.0);
        r.pairs("torqueRatioMap", flat, 1.0);
    } else {
        flat.push_back(0.0);
        flat.push_back(0.0);
    }
    if (!r.ok) return false;

    JValue t;
    if (!json_load_file(tire, t, err)) return false;
    Req q{t, tire, err};
    const double s2 = dof == 11 ? 2.0 : 1.0;  // axle factor of the 11-DOF loaders
    if (!tire_nr) {
        // setTireParamsJSON(TMeasyParam&), dof18.cpp:1218-1272 (dof11.cpp:1036-1086 doubles the
        // stiffnesses, damping, inertia, loads and forces); mu is not read (default 0.8)
        double v[35];
        const char* keys[] = {"jw", "rr", nullptr, "r0", "pn", "pnmax", "cx", "cy", "kt", "dx", "dy", "rdyncoPn",
                              "rdyncoP2n", "fzRdynco", "rdyncoCrit"};
        const bool axle[15] = {1, 1, 0, 0, 1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1};
        for (int i = 0; i < 15; ++i) v[i] = keys[i] ? q.num(keys[i]) * (axle[i] ? s2 : 1.0) : 0.8;
        for (int i = 0; i < 20; ++i) v[15 + i] = q.num(kCurveKeys[i]) * (kCurveAxle[i] ? s2 : 1.0);
        if (!q.ok) return false;
        tire_init(v[4], v[11], v[12], v[13], v[14]);
        flat.insert(flat.end(), v, v + 35);
        return dof == 24 ? read_suspension(sus, flat, err) : true;
    }
    // setTireParamsJSON(TMeasyNrParam&), dof18.cpp:1275-1461
    NrTire n;
    const JValue* hl = t.find("highLevelParams");
    if (hl && hl->kind == JValue::Bool && hl->b) {
        n.jw = q.num("jw") * s2; n.rr = q.num("rr") * s2; n.mu = q.num("mu") * s2; n.r0 = q.num("r0");
        n.rdyncoPn = q.num("rdyncoPn"); n.rdyncoP2n = q.num("rdyncoP2n");
        n.fzRdynco = q.num("fzRdynco") * s2; n.rdyncoCrit = q.num("rdyncoCrit") * s2;
        const JValue* vc = t.find("vCoulomb");
        const JValue* fb = t.find("frblendBegin");
        const JValue* fe = t.find("frblendEnd");
        n.vcoulomb = (vc && vc->kind == JValue::Number) ? vc->num : 1.0;
        n.frblend_begin = (fb && fb->kind == JValue::Number) ? fb->num : 1.0;
        n.frblend_end = (fe && fe->kind == JValue::Number) ? fe->num : 3.0;
        double width = q.num("width");
        double rimRadius = q.num("rimRadius");
        double p_li = 1.0, p_use = 1.0;
        double load;
        if (t.has("loadIndex")) {
            pressures(t, p_li, p_use);
            double li = q.num("loadIndex");
            if (q.ok && (li < 0 || li != std::floor(li))) { err = tire + ": loadIndex must be a non-negative integer"; return false; }
            n.li = li;
            load = tire_max_load((unsigned)li);
        } else {
            pressures(t, p_li, p_use);
            load = q.num("maximumBearingCapacity");
            n.bearingCapacity = load;
        }
        n.p_li = p_li;
        n.p_use = p_use;
        const JValue* vt = q.get("vehicleType", JValue::String);
        if (!q.ok) return false;
        const TireClass& k = vt->str == "Truck" ? kTruck80 : kPassCar70;
        guess_tire(k, load, width, (n.r0 - rimRadius) / width, 2 * rimRadius, p_li, p_use, n);
        // the 11-DOF guessers double the forces and slopes (dof11.cpp:1328-1380)
        for (int i = 0; i < 20; ++i) n.c[i] = n.c[i] * (kCurveAxle[i] ? s2 : 1.0);
    } else {
        nr_low_level(q, n, s2);
    }
    if (!q.ok) return false;
    tire_init(n.pn, n.rdyncoPn, n.rdyncoP2n, n.fzRdynco, n.rdyncoCrit);
    const double head[] = {n.jw, n.rr, n.mu, n.r0, n.width, n.rim_radius, n.pn, n.pnmax, n.kt,
                           n.rdyncoPn, n.rdyncoP2n, n.fzRdynco, n.rdyncoCrit};
    flat.insert(flat.end(), head, head + 13);
    flat.insert(flat.end(), n.c, n.c + 20);
    const double tail[] = {n.vcoulomb, n.frblend_begin, n.frblend_end, n.bearingCapacity, n.li, n.p_li, n.p_use};
    flat.insert(flat.end(), tail, tail + 7);
    return dof == 24 ? read_suspension(sus, flat, err) : true;
}

namespace {

struct Rd {
    const double* p;
    int len;
    int i = 0;
    bool ok = true;
    double next() {
        if (i >= len) { ok = false; return 0.0; }
        return p[i++];
    }
    bool map(Map& m, std::string& err, const char* name, bool required) {
        double nd = next();
        int n = (int)nd;
        if (!ok || nd != (double)n || n < 0) { err = std::string("bad size for map ") + name; ok = false; return false; }
        if (n > kMaxMap) { err = std::string("map ") + name + " has more than " + std::to_string(kMaxMap) + " entries"; ok = false; return false; }
        if (required && n == 0) { err = std::string("map ") + name + " is empty"; ok = false; return false; }
        m.n = n;
        m.pad = 0;
        for (int k = 0; k < n; ++k) {
            m.e[k].x = next();
            m.e[k].y = next();
        }
        for (int k = n; k < kMaxMap; ++k) m.e[k].x = m.e[k].y = 0.0;
        return ok;
    }
};

}  // namespace

namespace {

// Clamped piecewise-linear lookup with std::lower_bound's interval (utils.cpp:109-125).
double host_map_y(const Map& m, double x) {
    if (x <= m.e[0].x) return m.e[0].y;
    if (x >= m.e[m.n - 1].x) return m.e[m.n - 1].y;
    int r = 0;
    while (r < m.n && m.e[r].x < x) ++r;
    const double mbar = (x - m.e[r - 1].x) / (m.e[r].x - m.e[r - 1].x);
    return m.e[r - 1].y + mbar * (m.e[r].y - m.e[r - 1].y);
}

// Parameter-only subexpressions, written exactly as the reference writes them inside
// computeVehRHS (dof18.cpp:1019-1078), computeTireRHS / InterpQ / InterpL (utils.h:93-100) and
// computePowertrainRHS (dof18.cpp:799-844).
void derive18(PSet& p) {
    const TireP& t = p.tire;
    Derived& u = p.d;
    u.two_ab = 2. * (p.a + p.b);
    u.E2c = -p.muf * p.a + p.mur * p.b;
    u.Uc = -p.mur * p.b + p.muf * p.a;
    u.A1 = p.mur * p.b - p.muf * p.a;
    u.A1sq = u.A1 * u.A1;
    u.jxz2 = p.jxz * p.jxz;
    u.kro = p.krof + p.kror;
    u.bro = p.brof + p.bror;
    u.inv_jw = 1 / t.jw;
    u.inv_ci = 1. / p.crankInertia;
    u.qx1 = 2. * t.dfx0Pn - 0.5 * t.dfx0P2n;  u.qx2 = t.dfx0Pn - 0.5 * t.dfx0P2n;
    u.qfxm1 = 2. * t.fxmPn - 0.5 * t.fxmP2n;  u.qfxm2 = t.fxmPn - 0.5 * t.fxmP2n;
    u.qfym1 = 2. * t.fymPn - 0.5 * t.fymP2n;  u.qfym2 = t.fymPn - 0.5 * t.fymP2n;
    u.qfxs1 = 2. * t.fxsPn - 0.5 * t.fxsP2n;  u.qfxs2 = t.fxsPn - 0.5 * t.fxsP2n;
    u.qfys1 = 2. * t.fysPn - 0.5 * t.fysP2n;  u.qfys2 = t.fysPn - 0.5 * t.fysP2n;
    u.lsxm = t.sxmP2n - t.sxmPn;
    u.lsym = t.symP2n - t.symPn;
    u.lsxs = t.sxsP2n - t.sxsPn;
    u.lsys = t.sysP2n - t.sysPn;
    u.lrd = t.rdyncoP2n - t.rdyncoPn;
    u.delta_rear = p.nonLinearSteer ? host_map_y(p.steer, 0.0) : 0.0 * p.maxSteer;
    u.tan_delta_rear = std::tan(u.delta_rear);
    u.inv_pn = 1. / t.pn;
    u.inv_kt = 1. / t.kt;
    u.inv_mu = 1. / t.mu;
    u.inv_two_ab = 1. / u.two_ab;
    u.krof_cf = p.krof / p.cf;
    u.kror_cr = p.kror / p.cr;
    u.brof_cf = p.brof / p.cf;
    u.bror_cr = p.bror / p.cr;
    u.inv_cf = 1. / p.cf;
    u.inv_cr = 1. / p.cr;
}

}  // namespace

bool pset_from_flat(int dof, const double* flat, int len, int tire_nr, PSet& o, std::string& err) {
    o = PSet{};
    if (dof != 11 && dof != 18 && dof != 24) {
        err = "unsupported model: " + std::to_string(dof) + "-DOF";
        return false;
    }
    o.dof = dof;
    Rd r{flat, len};
    double* s11[] = {&o.a, &o.b, &o.h, &o.m, &o.jz, &o.muf, &o.mur};
    double* s18[] = {&o.a,   &o.b,    &o.h,    &o.m,    &o.jz,   &o.jx,   &o.jxz,  &o.cf,  &o.cr,
                     &o.muf, &o.mur,  &o.hrcf, &o.hrcr, &o.krof, &o.kror, &o.brof, &o.bror};
    double* s24[] = {&o.a, &o.b, &o.h, &o.m, &o.jz, &o.jx, &o.jy, &o.cf, &o.cr, &o.muf, &o.mur, &o.hrcf, &o.hrcr};
    double** s = dof == 11 ? s11 : (dof == 18 ? s18 : s24);
    const int ns = dof == 11 ? 7 : (dof == 18 ? 17 : 13);
    for (int i = 0; i < ns; ++i) *s[i] = r.next();
    o.maxSteer = r.next();
    o.crankInertia = r.next();
    o.maxBrakeTorque = r.next();
    o.nonLinearSteer = r.next() != 0;
    o.tcbool = r.next() != 0;
    o.throttleMod = r.next() != 0;
    o.driveType = r.next() != 0;
    o.whichWheels = r.next() != 0;
    double ngd = r.next();
    int ng = (int)ngd;
    if (!r.ok || ng < 1 || ng > kMaxGears || ngd != (double)ng) { err = "gear count must be 1.." + std::to_string(kMaxGears); return false; }
    o.ngears = ng;
    for (int i = 0; i < ng; ++i) o.gears[i] = r.next();
    Map shift;
    if (!r.map(shift, err, "shift", true)) return false;
    if (shift.n != ng) { err = "shift map needs one entry per gear"; return false; }
    for (int i = 0; i < ng; ++i) o.shift[i] = shift.e[i];
    if (!r.map(o.steer, err, "steer", o.nonLinearSteer != 0)) return false;
    if (!r.map(o.torque, err, "torque", true)) return false;
    if (!r.map(o.losses, err, "losses", true)) return false;
    if (!r.map(o.cfmap, err, "capacity factor", o.tcbool != 0)) return false;
    if (!r.map(o.trmap, err, "torque ratio", o.tcbool != 0)) return false;
    TireP& t = o.tire;
    t = TireP{};
    o.tire_nr = tire_nr ? 1 : 0;
    if (!tire_nr) {
        double* ts[] = {&t.jw, &t.rr, &t.mu, &t.r0, &t.pn, &t.pnmax, &t.cx, &t.cy, &t.kt, &t.dx, &t.dy, &t.rdyncoPn,
                        &t.rdyncoP2n, &t.fzRdynco, &t.rdyncoCrit, &t.dfx0Pn, &t.dfx0P2n, &t.fxmPn, &t.fxmP2n,
                        &t.fxsPn, &t.fxsP2n, &t.sxmPn, &t.sxmP2n, &t.sxsPn, &t.sxsP2n, &t.dfy0Pn, &t.dfy0P2n,
                        &t.fymPn, &t.fymP2n, &t.fysPn, &t.fysP2n, &t.symPn, &t.symP2n, &t.sysPn, &t.sysP2n};
        for (double* p : ts) *p = r.next();
    } else {
        double* ts[] = {&t.jw, &t.rr, &t.mu, &t.r0, &t.width, &t.rim_radius, &t.pn, &t.pnmax, &t.kt, &t.rdyncoPn,
                        &t.rdyncoP2n, &t.fzRdynco, &t.rdyncoCrit, &t.dfx0Pn, &t.dfx0P2n, &t.fxmPn, &t.fxmP2n,
                        &t.fxsPn, &t.fxsP2n, &t.sxmPn, &t.sxmP2n, &t.sxsPn, &t.sxsP2n, &t.dfy0Pn, &t.dfy0P2n,
                        &t.fymPn, &t.fymP2n, &t.fysPn, &t.fysP2n, &t.symPn, &t.symP2n, &t.sysPn, &t.sysP2n,
                        &t.vcoulomb, &t.frblend_begin, &t.frblend_end, &t.bearingCapacity, &t.li, &t.p_li, &t.p_use};
        for (double* p : ts) *p = r.next();
    }
    if (dof == 24) {
        o.ks = r.next();
        o.bs = r.next();
    }
    if (!r.ok) { err = "flat parameter vector too short"; return false; }
    if (r.i != len) { err = "flat parameter vector has " + std::to_string(len - r.i) + " trailing values"; return false; }
    derive18(o);
    return true;
}

bool flat_slots(int dof, const double* flat, int len, int tire_nr, FlatSlots& o, std::string& err) {
    PSet check;
    if (!pset_from_flat(dof, flat, len, tire_nr, check, err)) return false;  // validates the layout
    const int nscal = dof == 11 ? 10 : (dof == 18 ? 20 : 16);
    o.m = 3;
    o.jz = 4;
    o.jx = dof == 11 ? -1 : 5;
    o.max_brake = nscal - 1;
    int i = nscal + 5;  // after the five flags: the gear count
    i += 1 + (int)flat[i];
    for (int map = 0; map < 6; ++map) {  // shift, steer, torque, losses, capacity factor, torque ratio
        const int n = (int)flat[i];
        if (map == 2) {
            o.torque = i + 1;
            o.torque_n = n;
        }
        i += 1 + 2 * n;
    }
    o.dfy0Pn = i + (tire_nr ? 23 : 25);
    o.dfy0P2n = o.dfy0Pn + 1;
    return true;
}

bool load_driver_file(const std::string& path, std::vector<DrvRow>& rows, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open driver file " + path;
        return false;
    }
    rows.clear();
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream iss(line);
        iss.imbue(std::locale::classic());  // '.' decimals whatever the process's locale
        DrvRow r;
        iss >> r.t >> r.steering >> r.throttle >> r.braking;
        if (iss.fail()) break;
        rows.push_back(r);
    }
    if (rows.empty()) {
        err = "driver file " + path + " holds no rows";
        return false;
    }
    return true;
}

void init_tire_sus24(const PSet& p, double m, int tire_nr, double* rec) {
    const double g = 9.81;
    const int tl = tire_nr ? 19 : 23;         // tire block length
    const int o_xt = tire_nr ? 2 : 6, o_fz = tire_nr ? 7 : 11;
    double* sus = rec + 4 * tl;               // strut blocks: xs xsi dxs ls lsi ... wu at 13
    const double w = rec[4 * tl + 4 * 24 + 5];  // vehicle heave velocity
    double fz[4];
    fz[0] = fz[1] = ((m * g * p.b) / (2 * (p.a + p.b)) + p.muf * g);
    fz[2] = fz[3] = ((m * g * p.a) / (2 * (p.a + p.b)) + p.mur * g);
    double xsi[4];
    xsi[0] = xsi[1] = (m * g * p.b) / (2 * (p.a + p.b) * p.ks);
    xsi[2] = xsi[3] = (m * g * p.a) / (2 * (p.a + p.b) * p.ks);
    for (int i = 0; i < 4; ++i) {
        const double xt = fz[i] / p.tire.kt;
        rec[i * tl + o_fz] = fz[i];
        rec[i * tl + o_xt] = xt;
        double* q = sus + 24 * i;
        q[0] = xsi[i];                   // xs
        q[1] = xsi[i];                   // xsi
        q[4] = p.h - (p.tire.r0 - xt);   // lsi
        q[3] = q[4];                     // ls
        q[13] = w;                       // wu
    }
}

}  // namespace lfdm
