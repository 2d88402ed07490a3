// A small test harness in the style of the reference's Boost tests:
// TEST_CASE registers a function, CHECK / CHECK_CLOSE (percent tolerance, as
// BOOST_CHECK_CLOSE) / CHECK_THROW record failures, main() runs every case
// (or those named on the command line) and prints OK when all pass.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

namespace check {
struct Case {
    const char *name;
    std::function<void()> fn;
};
inline std::vector<Case> &cases()
{
    static std::vector<Case> c;
    return c;
}
inline int &failures()
{
    static int f = 0;
    return f;
}
struct Reg {
    Reg(const char *n, std::function<void()> f) { cases().push_back({n, std::move(f)}); }
};
inline void fail(const char *file, int line, const std::string &msg)
{
    ++failures();
    std::printf("%s:%d: FAILED %s\n", file, line, msg.c_str());
}
inline bool close_pct(double a, double b, double pct)
{
    if (a == b)
        return true;
    const double d = std::fabs(a - b);
    return d <= pct / 100.0 * std::fabs(a) && d <= pct / 100.0 * std::fabs(b);
}
} // namespace check

#define CHECK_CAT2(a, b) a##b
#define CHECK_CAT(a, b) CHECK_CAT2(a, b)
#define TEST_CASE(name)                                                                                       \
    static void name();                                                                                       \
    static check::Reg CHECK_CAT(reg_, name)(#name, name);                                                     \
    static void name()
#define CHECK(c)                                                                                              \
    do {                                                                                                      \
        if (!(c))                                                                                             \
            check::fail(__FILE__, __LINE__, #c);                                                              \
    } while (0)
#define CHECK_CLOSE(a, b, pct)                                                                                \
    do {                                                                                                      \
        const double a_ = (a), b_ = (b);                                                                      \
        if (!check::close_pct(a_, b_, (pct)))                                                                 \
            check::fail(__FILE__, __LINE__,                                                                   \
                        std::string(#a " ~ " #b ": ") + std::to_string(a_) + " vs " + std::to_string(b_));     \
    } while (0)
#define CHECK_THROW(expr, exc)                                                                                \
    do {                                                                                                      \
        bool thrown_ = false;                                                                                 \
        try {                                                                                                 \
            (void)(expr);                                                                                     \
        } catch (const exc &) {                                                                               \
            thrown_ = true;                                                                                   \
        } catch (...) {                                                                                       \
        }                                                                                                     \
        if (!thrown_)                                                                                         \
            check::fail(__FILE__, __LINE__, "expected " #exc " from " #expr);                                 \
    } while (0)

#define CHECK_MAIN()                                                                                          \
    int main(int argc, char **argv)                                                                           \
    {                                                                                                         \
        int run = 0;                                                                                          \
        for (auto &c : check::cases()) {                                                                      \
            bool sel = argc < 2;                                                                              \
            for (int i = 1; i < argc; ++i)                                                                    \
                sel |= std::strcmp(argv[i], c.name) == 0;                                                     \
            if (!sel)                                                                                         \
                continue;                                                                                     \
            const int before = check::failures();                                                             \
            try {                                                                                             \
                c.fn();                                                                                       \
            } catch (const std::exception &e) {                                                               \
                check::fail(c.name, 0, std::string("exception: ") + e.what());                                \
            }                                                                                                 \
            std::printf("%-48s %s\n", c.name, check::failures() == before ? "ok" : "FAILED");                 \
            std::fflush(stdout);                                                                              \
            ++run;                                                                                            \
        }                                                                                                     \
        std::printf("%d cases, %d failures\n", run, check::failures());                                      \
        if (check::failures() == 0)                                                                           \
            std::printf("OK\n");                                                                              \
        return check::failures() == 0 ? 0 : 1;                                                                \
    }
