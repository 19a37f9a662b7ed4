// The random values a driver of the reference draws for its initial circles
// (mainIsoElpsPack_WCCM.cpp:46-66): std::default_random_engine seeded with 0
// and std::uniform_real_distribution<double>, consumed through a constructor
// call of the same shape, so that g++'s argument evaluation order applies.
//   rng_probe <n> <aspect_ratio>   prints "alpha radius x y orientation" lines
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

namespace {
struct Vec {
    double x, y;
    Vec(double a, double b) : x(a), y(b) {}
};
struct Circle {
    double r, x, y, t;
    Circle(int, double radius, Vec pos, Vec, double rot, double, const void*)
        : r(radius), x(pos.x), y(pos.y), t(rot) {}
};
}  // namespace

int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 10;
    const double ar = argc > 2 ? atof(argv[2]) : 2.0;
    std::default_random_engine engine(0);
    std::uniform_real_distribution<double> real_r(2.0 / 3.0, 4.0 / 3.0);
    std::uniform_real_distribution<double> real_x(-100.0, 100.0);
    std::uniform_real_distribution<double> real_y(-100.0, 100.0);
    std::uniform_real_distribution<double> real_t(0.0, M_PI * 2.0);
    std::uniform_real_distribution<double> real_a(1.01, ar);
    const double f = std::sqrt(ar);
    for (int i = 0; i < n; ++i) {
        const double a = real_a(engine);
        const Circle c(i, real_r(engine) * f, Vec(real_x(engine), real_y(engine)), Vec(0.0, 0.0),
                       real_t(engine), 0.0, nullptr);
        printf("%.17g %.17g %.17g %.17g %.17g\n", a, c.r, c.x, c.y, c.t);
    }
    return 0;
}
