// prc_cli.cpp -- the `prc ISOTOPE NX NY NZ DX DY DZ [options]` executable.
//
// Drop-in for the reference CLI (prc.cu:452-722): same argv grammar, defaults, validation,
// stderr messages, exit codes, raw float32 I/O and output naming. Every check runs before
// the first call that touches the GPU, so rejected runs work with no device visible. The
// numerical work goes through the C ABI of libprc_amd.so (include/prc.h).
#include <algorithm>
#include <cerrno>
#include <cstdarg>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "prc.h"

namespace {

constexpr float kMinDensity = 1.0e-3f;  // prc.cu:102

// A rejected command line or input: the message goes to stderr (optionally followed by the usage
// text) and the program exits with EXIT_FAILURE, before anything touches the GPU.
struct CliError {
    std::string message;
    bool usage = false;
};

[[noreturn]] void reject(bool usage, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw CliError{buf, usage};
}

// Settings of one run with the reference's defaults (prc.cu:452-459).
struct Options {
    std::string isotope, mode = "correct", input = "ACTIVITY.raw", density = "DENSITY.raw", output;
    std::string precision = "fp32";
    int nx = 0, ny = 0, nz = 0;
    float dx = 0, dy = 0, dz = 0;
    int rays = 128, iterations = 20, save_every = 0;
    float beta = 0.2f, gamma = 20.0f, support_rho = 0.05f, air_cap_mm = 50.0f;
    bool reproject = false, cache_near = true, exact_range = false, accelerate = true;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void print_usage(const char* prog, FILE* stream = stderr) {
    fprintf(stream,
            "Usage: %s ISOTOPE NX NY NZ DX DY DZ [options]\n"
            "  --mode correct|forward|adjoint   (default correct)\n"
            "  --input FILE      blurred (correct) or source (forward/adjoint) image, default ACTIVITY.raw\n"
            "  --density FILE    density map in g/cm3, default DENSITY.raw\n"
            "  --output FILE     default ACTIVITY_PRC_V2.raw (correct), ACTIVITY_FWD.raw, ACTIVITY_ADJ.raw\n"
            "  --rays N          even far-field rays per voxel (default 128)\n"
            "  --iterations N    accelerated MAP-EM iterations (default 20)\n"
            "  --beta B          relative-difference prior strength (default 0.2; 0 = no prior)\n"
            "                    Add --no-accel for plain Richardson-Lucy.\n"
            "  --gamma G         RDP edge preservation, larger keeps sharper edges (default 20)\n"
            "  --support-rho R   minimum density of voxels that can hold activity (default 0.05)\n"
            "  --air-cap-mm L    a positron that crossed L mm of air has escaped (default 50)\n"
            "  --save-every K    also write the estimate every K iterations\n"
            "  --reproject       also write A f of the result (_REPROJ.raw)\n"
            "  --exact-range     evaluate the survival function analytically, not from the table (validation)\n"
            "  --no-cache-near   store only the near block and retrace the far rays in each pass\n"
            "                    (less device memory, same operator)\n"
            "  --no-accel        plain MAP-EM without Biggs-Andrews extrapolation (validation)\n"
            "  --precision P     fp32 (default) or mixed: FP16 stored weights, FP32 arithmetic\n"
            "Isotopes:",
            prog);
    for (int i = 0; i < prc_isotope_count(); ++i) fprintf(stream, " %s", prc_isotope_name(i));
    fprintf(stream, "\n");
}

// Strict number parsing: the whole text must be one finite number in range (prc.cu:486-506 states
// the same rules and messages).
int to_int(const char* text, const char* what) {
    char* end = nullptr;
    errno = 0;
    const long v = strtol(text, &end, 10);
    const bool whole = end != text && *end == '\0';
    if (errno != 0 || !whole || v < INT_MIN || v > INT_MAX) reject(false, "Invalid integer for %s: %s", what, text);
    return (int)v;
}

float to_float(const char* text, const char* what) {
    char* end = nullptr;
    errno = 0;
    const float v = strtof(text, &end);
    const bool whole = end != text && *end == '\0';
    if (errno != 0 || !whole || !std::isfinite(v)) reject(false, "Invalid finite number for %s: %s", what, text);
    return v;
}

// The options the command line accepts: a switch sets a flag, the others take the next argument.
struct OptionSpec {
    const char* name;
    enum Kind { kSwitch, kInt, kFloat, kText } kind;
    void* target;
    bool value = true;  // what a switch stores
};

std::vector<OptionSpec> option_table(Options& o) {
    return {
        {"--mode", OptionSpec::kText, &o.mode},
        {"--input", OptionSpec::kText, &o.input},
        {"--density", OptionSpec::kText, &o.density},
        {"--output", OptionSpec::kText, &o.output},
        {"--precision", OptionSpec::kText, &o.precision},
        {"--rays", OptionSpec::kInt, &o.rays},
        {"--iterations", OptionSpec::kInt, &o.iterations},
        {"--save-every", OptionSpec::kInt, &o.save_every},
        {"--beta", OptionSpec::kFloat, &o.beta},
        {"--gamma", OptionSpec::kFloat, &o.gamma},
        {"--support-rho", OptionSpec::kFloat, &o.support_rho},
        {"--air-cap-mm", OptionSpec::kFloat, &o.air_cap_mm},
        {"--reproject", OptionSpec::kSwitch, &o.reproject, true},
        {"--no-cache-near", OptionSpec::kSwitch, &o.cache_near, false},
        {"--exact-range", OptionSpec::kSwitch, &o.exact_range, true},
        {"--no-accel", OptionSpec::kSwitch, &o.accelerate, false},
    };
}

// ISOTOPE NX NY NZ DX DY DZ [options], then the range checks (prc.cu:508-555 has the same grammar,
// checks and messages: they are the drop-in contract).
Options parse(int argc, char** argv) {
    if (argc < 8) reject(true, "");
    Options o;
    o.isotope = argv[1];
    int* dims[3] = {&o.nx, &o.ny, &o.nz};
    float* sizes[3] = {&o.dx, &o.dy, &o.dz};
    const char* dim_names[3] = {"NX", "NY", "NZ"};
    const char* size_names[3] = {"DX", "DY", "DZ"};
    for (int a = 0; a < 3; ++a) *dims[a] = to_int(argv[2 + a], dim_names[a]);
    for (int a = 0; a < 3; ++a) *sizes[a] = to_float(argv[5 + a], size_names[a]);
    const std::vector<OptionSpec> table = option_table(o);
    for (int i = 8; i < argc; ++i) {
        const char* key = argv[i];
        const OptionSpec* spec = nullptr;
        for (const OptionSpec& s : table)
            if (!strcmp(s.name, key)) spec = &s;
        if (!spec) reject(true, "Unknown option %s", key);
        if (spec->kind == OptionSpec::kSwitch) {
            *(bool*)spec->target = spec->value;
            continue;
        }
        if (i + 1 >= argc) reject(false, "Missing value for %s", key);
        const char* text = argv[++i];
        switch (spec->kind) {
            case OptionSpec::kInt: *(int*)spec->target = to_int(text, key); break;
            case OptionSpec::kFloat: *(float*)spec->target = to_float(text, key); break;
            default: *(std::string*)spec->target = text; break;
        }
    }
    const bool known_mode = o.mode == "correct" || o.mode == "forward" || o.mode == "adjoint";
    const bool ok = o.nx > 0 && o.ny > 0 && o.nz > 0 && o.dx > 0 && o.dy > 0 && o.dz > 0 && o.rays >= 2 &&
                    o.rays % 2 == 0 && o.rays <= INT_MAX - 31 && o.iterations >= 0 && o.iterations < INT_MAX &&
                    o.beta >= 0 && o.gamma >= 0 && o.support_rho >= 0 && o.air_cap_mm > 0 && o.save_every >= 0 &&
                    known_mode && (o.precision == "fp32" || o.precision == "mixed");
    if (!ok) reject(false, "Invalid geometry or solver parameters");
    // Voxel indices are signed 32-bit: the products are checked before they are formed.
    if (o.nx > INT_MAX / o.ny || o.nx * o.ny > INT_MAX / o.nz) reject(false, "Volume exceeds the INT_MAX voxel index limit");
    if (o.output.empty()) {
        const char* by_mode = o.mode == "correct" ? "ACTIVITY_PRC_V2.raw" : o.mode == "forward" ? "ACTIVITY_FWD.raw"
                                                                                               : "ACTIVITY_ADJ.raw";
        o.output = by_mode;
    }
    return o;
}

// `path` with `tag` before the extension of its file name; a dot inside a directory name is not an
// extension (the reference's naming, prc.cu:557-562).
std::string tagged(const std::string& path, const std::string& tag) {
    const size_t name = path.find_last_of("/\\") == std::string::npos ? 0 : path.find_last_of("/\\") + 1;
    const size_t ext = path.rfind('.');
    if (ext == std::string::npos || ext <= name) return path + tag;
    return path.substr(0, ext) + tag + path.substr(ext);
}

// A headerless float32 volume of exactly n values (prc.cu:339-347: the size must match).
std::vector<float> load_volume(const std::string& path, size_t n) {
    std::vector<float> v(n);
    FILE* fp = fopen(path.c_str(), "rb");
    bool ok = fp && fread(v.data(), sizeof(float), n, fp) == n && fgetc(fp) == EOF && !ferror(fp);
    if (fp) fclose(fp);
    if (!ok) reject(false, "Expected exactly %zu float32 values in %s", n, path.c_str());
    return v;
}

// Writes a volume; a non-finite value rejects the output before the file is opened
// (prc.cu:349-360).
void write_volume(const std::string& path, const std::vector<float>& v) {
    if (std::any_of(v.begin(), v.end(), [](float x) { return !std::isfinite(x); }))
        reject(false, "Non-finite output: %s", path.c_str());
    FILE* fp = fopen(path.c_str(), "wb");
    const bool ok = fp && fwrite(v.data(), sizeof(float), v.size(), fp) == v.size();
    if (!fp || fclose(fp) != 0 || !ok) reject(false, "Cannot write %s", path.c_str());
}

void check(int rc) {
    if (rc != PRC_OK) reject(false, "prc: %s", prc_last_error());
}

double sum_of(const std::vector<float>& v) {
    double s = 0.0;
    for (float x : v) s += x;
    return s;
}

int run(int argc, char** argv) {
    const Options o = parse(argc, argv);
    const int iso = prc_isotope_find(o.isotope.c_str());
    if (iso < 0) reject(true, "");
    const float cutoff = prc_range_cutoff(iso, 0.9995f);  // prc.cu:587
    const size_t n = (size_t)o.nx * o.ny * o.nz;

    const double t0 = now_s();
    std::vector<float> rho = load_volume(o.density, n), img = load_volume(o.input, n);
    size_t n_active = 0;
    double input_sum = 0.0, support_sum = 0.0;
    for (size_t i = 0; i < n; ++i) {  // preprocessing, prc.cu:598-613
        if (!std::isfinite(rho[i]) || !std::isfinite(img[i])) reject(false, "Non-finite density or activity at voxel %zu", i);
        rho[i] = std::max(rho[i], kMinDensity);
        img[i] = std::max(img[i], 0.0f);
        input_sum += img[i];
        const bool tissue = rho[i] >= o.support_rho;
        n_active += tissue;
        support_sum += tissue ? img[i] : 0.0f;
        if (!tissue) img[i] = 0.0f;
    }
    printf("isotope %s, grid %dx%dx%d, voxel %.3fx%.3fx%.3f mm, cutoff %.3f mm (water), air cap %.1f mm\n",
           o.isotope.c_str(), o.nx, o.ny, o.nz, o.dx, o.dy, o.dz, cutoff, o.air_cap_mm);
    printf("tissue voxels %zu (%.1f%%), input counts in tissue %.6g of %.6g (%.2f%%)\n", n_active,
           100.0 * n_active / n, support_sum, input_sum, input_sum > 0 ? 100.0 * support_sum / input_sum : 0.0);
    if (support_sum == 0.0) reject(false, "No activity in voxels with density >= %g: nothing to do", o.support_rho);

    // First call that touches the GPU.
    prc_geometry g{};
    g.nx = o.nx, g.ny = o.ny, g.nz = o.nz, g.sx = o.dx, g.sy = o.dy, g.sz = o.dz;
    g.isotope = iso, g.n_rays = o.rays, g.support_rho = o.support_rho, g.air_cap_mm = o.air_cap_mm;
    g.precision = o.precision == "mixed" ? PRC_PRECISION_MIXED : PRC_PRECISION_FP32;
    g.flags = (o.exact_range ? PRC_FLAG_EXACT_RANGE : 0) | (o.cache_near ? 0 : PRC_FLAG_MARCH_PASSES);
    double t = now_s();
    prc_projector* A = nullptr;
    check(prc_projector_create(&g, rho.data(), nullptr, &A));
    prc_stats st{};
    check(prc_projector_stats(A, &st));
    printf("system matrix: %lld stored weights (%.0f MiB, %s%s)\n", (long long)st.weight_entries,
           st.weight_bytes / 1048576.0, o.precision.c_str(), st.far_marched ? ", far rays marched in each pass" : "");
    printf("normalisation (%d near-block + %d far-field rays): %.2f s\n", PRC_NEAR_RAYS, o.rays, now_s() - t);

    if (o.mode != "correct") {
        t = now_s();
        std::vector<float> result(n, 0.0f);
        if (o.mode == "forward") check(prc_projector_forward_host(A, img.data(), result.data()));
        else check(prc_projector_adjoint_host(A, img.data(), result.data()));
        printf("%s: %.3f s\n", o.mode.c_str(), now_s() - t);
        write_volume(o.output, result);
        printf("sum in %.6g, sum out %.6g, total %.2f s\n", support_sum, sum_of(result), now_s() - t0);
        prc_projector_destroy(A);
        return EXIT_SUCCESS;
    }

    printf("MAP-EM: %d iterations, beta %.4g, gamma %.3g\n", o.iterations, o.beta, o.gamma);
    prc_solve_params p{o.beta, o.gamma, o.accelerate ? 1 : 0};
    prc_solver* S = nullptr;
    check(prc_solver_create(A, img.data(), &p, &S));
    std::vector<float>& out = img;  // the input buffer holds the outputs
    auto reports = [&](int it) { return it == 1 || it % 10 == 0 || it == o.iterations; };  // prc.cu:701-704
    auto saves = [&](int it) { return o.save_every > 0 && it % o.save_every == 0 && it < o.iterations; };
    t = now_s();
    // Iterations run on the device without synchronisation up to the next one that prints or saves.
    for (int done = 0; done < o.iterations;) {
        int next = done + 1;
        while (next < o.iterations && !reports(next) && !saves(next)) ++next;
        check(prc_solver_iterate(S, next - done));
        done = next;
        if (reports(done)) {
            check(prc_solver_sync(S));
            printf("iteration %d/%d: %.2f s\n", done, o.iterations, now_s() - t);
        }
        if (saves(done)) {
            char tag[32];
            snprintf(tag, sizeof(tag), "_it%03d", done);
            check(prc_solver_estimate(S, out.data()));
            write_volume(tagged(o.output, tag), out);
        }
    }
    check(prc_solver_sync(S));
    const double t_iter = now_s() - t;
    check(prc_solver_estimate(S, out.data()));
    write_volume(o.output, out);
    const double corrected_sum = sum_of(out);
    if (o.reproject) {
        check(prc_solver_reproject(S, out.data()));
        write_volume(tagged(o.output, "_REPROJ"), out);
    }
    printf("sum blurred (tissue) %.6g, sum corrected %.6g\n", support_sum, corrected_sum);
    printf("iterations %.2f s (%.2f s each), total %.2f s\n", t_iter, o.iterations ? t_iter / o.iterations : 0.0,
           now_s() - t0);
    prc_solver_destroy(S);
    prc_projector_destroy(A);
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 2 && (!strcmp(argv[1], "--help") || !strcmp(argv[1], "-h"))) {
        print_usage(argv[0], stdout);
        return EXIT_SUCCESS;
    }
    try {
        return run(argc, argv);
    } catch (const CliError& e) {
        fflush(stdout);
        if (!e.message.empty()) fprintf(stderr, "%s\n", e.message.c_str());
        if (e.usage) print_usage(argv[0]);
        return EXIT_FAILURE;
    }
}
