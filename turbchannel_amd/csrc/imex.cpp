// Butcher tables of the IMEX Runge-Kutta methods of the collocated SRK stepper
// (collocated_SRK/imex.cpp:48-239, collocated_SRK/imex.h:26-42). The
// coefficients are the published ones:
//   ARS(1,2,1), ARS(2,3,2), ARS(3,4,3): Ascher, Ruuth, Spiteri, Appl. Numer.
//     Math. 25 (1997);
//   MARS(3,4,3), MARK3(2)4L[2]SA: Boscarino, Russo, SIAM J. Sci. Comput. (2007);
//   ARK3(2)4L[2]SA, ARK4(3)6L[2]SA: Kennedy, Carpenter, Appl. Numer. Math. 44 (2003);
//   BHR(5,5,3): Boscarino, Appl. Numer. Math. 59 (2009).
// Table layout: (s+1) rows of s entries, row i < s holding a_ij, row s the
// weights b_j; E explicit, I implicit.
#include <cmath>
#include <cstring>
#include <vector>

#include "tc_internal.h"

namespace tc {

namespace {

// gamma of the L-stable three-stage DIRK family: middle root of 6x^3 - 18x^2 + 9x - 1 = 0
constexpr double kGammaL = 0.435866521508458999416019451193556842529;

void set(ImexTable& t, int type, int s, std::initializer_list<double> E, std::initializer_list<double> I) {
    t.type = type;
    t.s = s;
    t.E.assign(E);
    t.I.assign(I);
}

}  // namespace

bool imex_table(int method, ImexTable& t) {
    t = ImexTable();
    switch (method) {
        case TC_IMEX_ARS_121:  // implicit-explicit Euler
            set(t, TC_IMEX_TYPE_ARS, 2, {0., 0., 1., 0., 0., 1.}, {0., 0., 0., 1., 0., 1.});
            break;
        case TC_IMEX_ARS_232: {
            const double g = 1. - 0.5 * std::sqrt(2.), dl = -2. * std::sqrt(2.) / 3.;
            set(t, TC_IMEX_TYPE_ARS, 3, {0., 0., 0., g, 0., 0., dl, 1. - dl, 0., 0., 1. - g, g},
                {0., 0., 0., 0., g, 0., 0., 1. - g, g, 0., 1. - g, g});
            break;
        }
        case TC_IMEX_ARS_343: {
            const double g = kGammaL;
            const double b1 = -1.5 * g * g + 4 * g - 0.25, b2 = 1.5 * g * g - 5. * g + 1.25;
            const double a42 = 0.5529291479, a43 = a42;
            const double p = 1. - 4.5 * g + 1.5 * g * g, q = 2.75 - 10.5 * g + 3.75 * g * g;
            const double a31 = p * a42 + q * a43 - 3.5 + 13. * g - 4.5 * g * g;
            const double a32 = -p * a42 - q * a43 + 4. - 12.5 * g + 4.5 * g * g;
            set(t, TC_IMEX_TYPE_ARS, 4,
                {0., 0., 0., 0., g, 0., 0., 0., a31, a32, 0., 0., 1. - a42 - a43, a42, a43, 0., 0., b1, b2, g},
                {0., 0., 0., 0., 0., g, 0., 0., 0., (1. - g) / 2., g, 0., 0., b1, b2, g, 0., b1, b2, g});
            break;
        }
        case TC_IMEX_MARS_343: {
            const double g = kGammaL, b2 = 1.20849664917601;
            set(t, TC_IMEX_TYPE_ARS, 4,
                {0., 0., 0., 0., g, 0., 0., 0., 0.535396540307354, 0.182536720446875, 0., 0., 0.63041255815287,
                 -0.83193390106308, 1.20152134291021, 0., 0., b2, 1. - g - b2, g},
                {0., 0., 0., 0., 0., g, 0., 0., 0., 0.28206673924577, g, 0., 0., b2, 1. - g - b2, g, 0., b2,
                 1. - g - b2, g});
            break;
        }
        case TC_IMEX_ARK3: {
            const double g = kGammaL;
            const double b1 = 1471266399579. / 7840856788654., b2 = -4482444167858. / 7529755066697.;
            const double b3 = 1. - g - b1 - b2;
            set(t, TC_IMEX_TYPE_CK, 4,
                {0., 0., 0., 0., 2 * g, 0., 0., 0., 5535828885825. / 10492691773637., 788022342437. / 10882634858940.,
                 0., 0., 6485989280629. / 16251701735622., -4246266847089. / 9704473918619.,
                 10755448449292. / 10357097424841., 0., b1, b2, b3, g},
                {0., 0., 0., 0., g, g, 0., 0., 2746238789719. / 10658868560708., -640167445237. / 6845629431997., g,
                 0., b1, b2, b3, g, b1, b2, b3, g});
            break;
        }
        case TC_IMEX_ARK4: {
            const double b[6] = {82889. / 524892., 0., 15625. / 83664., 69875. / 102672., -2260. / 8211., 0.25};
            set(t, TC_IMEX_TYPE_CK, 6,
                {0., 0., 0., 0., 0., 0.,  //
                 0.5, 0., 0., 0., 0., 0.,  //
                 13861. / 62500., 6889. / 62500., 0., 0., 0., 0.,  //
                 -116923316275. / 2393684061468., -2731218467317. / 15368042101831.,
                 9408046702089. / 11113171139209., 0., 0., 0.,  //
                 -451086348788. / 2902428689909., -2682348792572. / 7519795681897.,
                 12662868775082. / 11960479115383., 3355817975965. / 11060851509271., 0., 0.,  //
                 647845179188. / 3216320057751., 73281519250. / 8382639484533., 552539513391. / 3454668386233.,
                 3354512671639. / 8306763924573., 4040. / 17871., 0.,  //
                 b[0], b[1], b[2], b[3], b[4], b[5]},
                {0., 0., 0., 0., 0., 0.,  //
                 0.25, 0.25, 0., 0., 0., 0.,  //
                 8611. / 62500., -1743. / 31250., 0.25, 0., 0., 0.,  //
                 5012029. / 34652500., -654441. / 2922500., 174375. / 388108., 0.25, 0., 0.,  //
                 15267082809. / 155376265600., -71443401. / 120774400., 730878875. / 902184768.,
                 2285395. / 8070912., 0.25, 0.,  //
                 b[0], b[1], b[2], b[3], b[4], b[5],  //
                 b[0], b[1], b[2], b[3], b[4], b[5]});
            break;
        }
        case TC_IMEX_MARK3: {
            const double g = kGammaL, b1 = 0.60424832458800;
            set(t, TC_IMEX_TYPE_CK, 4,
                {0., 0., 0., 0., 2 * g, 0., 0., 0., -3.06478674186224, 1.46604002506519, 0., 0., 0.21444560762133,
                 0.71075364965269, 0.07480074272597, 0., b1, 0., 1. - g - b1, g},
                {0., 0., 0., 0., g, g, 0., 0., -4.30002662176923, 2.26541338346372, g, 0., b1, 0., 1. - g - b1, g,
                 b1, 0., 1. - g - b1, g});
            break;
        }
        case TC_IMEX_BHR_553: {
            t.type = TC_IMEX_TYPE_CK;
            t.s = 5;
            t.E.assign(30, 0.0);
            t.I.assign(30, 0.0);
            const double g = 424782. / 974569.;
            double* I = t.I.data();
            double* E = t.E.data();
            I[5] = I[6] = g;
            I[10] = g;
            I[11] = -31733082319927313. / 455705377221960889379854647102.;
            I[12] = g;
            I[15] = -3012378541084922027361996761794919360516301377809610. /
                    45123394056585269977907753045030512597955897345819349.;
            I[16] = -62865589297807153294268. / 102559673441610672305587327019095047.;
            I[17] = 418769796920855299603146267001414900945214277000. / 212454360385257708555954598099874818603217167139.;
            I[18] = g;
            I[20] = 487698502336740678603511. / 1181159636928185920260208.;
            I[22] = 302987763081184622639300143137943089. / 1535359944203293318639180129368156500.;
            I[23] = -105235928335100616072938218863. / 2282554452064661756575727198000.;
            I[24] = g;
            E[5] = 2. * g;
            E[10] = g;
            E[11] = g;
            E[15] = -475883375220285986033264. / 594112726933437845704163.;
            E[17] = 1866233449822026827708736. / 594112726933437845704163.;
            E[20] = 62828845818073169585635881686091391737610308247. / 176112910684412105319781630311686343715753056000.;
            E[21] = -I[22];
            E[22] = 262315887293043739337088563996093207. / 297427554730376353252081786906492000.;
            E[23] = -987618231894176581438124717087. / 23877337660202969319526901856000.;
            for (int i = 0; i < 5; i++) E[25 + i] = I[25 + i] = I[20 + i];
            break;
        }
        default:
            return false;
    }
    // d_j of the methods of type CK (collocated_SRK/imex.cpp:34-44): d_0 = 1,
    // d_j = -sum_{k<j} a_jk d_k / a_ss with a_ss the last diagonal entry of I
    const int s = t.s;
    const double ass = t.I[s * s - 1];
    t.d.assign(s, 0.0);
    t.d[0] = 1.;
    for (int j = 1; j < s; j++)
        for (int k = 0; k < j; k++) t.d[j] -= t.I[j * s + k] * t.d[k] / ass;
    t.alpha_tau_max[0] = t.alpha_tau_max[1] = 2.;
    return true;
}

}  // namespace tc

extern "C" int tc_imex_table(int method, int* nstages, int* type, double* be, double* bi, double* d,
                             double* alpha_tau_max) {
    tc::ImexTable t;
    if (!tc::imex_table(method, t)) return TC_ERR_ARG;
    if (nstages) *nstages = t.s;
    if (type) *type = t.type;
    if (be) std::memcpy(be, t.E.data(), t.E.size() * sizeof(double));
    if (bi) std::memcpy(bi, t.I.data(), t.I.size() * sizeof(double));
    if (d) std::memcpy(d, t.d.data(), t.d.size() * sizeof(double));
    if (alpha_tau_max) std::memcpy(alpha_tau_max, t.alpha_tau_max, 2 * sizeof(double));
    return TC_OK;
}
