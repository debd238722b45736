// prc_cli.cpp -- the `prc ISOTOPE NX NY NZ DX DY DZ [options]` executable.
//
// Drop-in for the reference CLI (prc.cu:452-722): same argv grammar, defaults, validation,
// stderr messages, exit codes, raw float32 I/O and output naming. Every check runs before
// the first call that touches the GPU, so rejected runs work with no device visible. The
// numerical work goes through the C ABI of libprc_amd.so (include/prc.h).
#include <cerrno>
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

struct Options {  // prc.cu:452-459
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

void print_usage(const char* prog, FILE* stream = stderr) {  // prc.cu:461-484
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
            "                    (less device memory, same result)\n"
            "  --no-accel        plain MAP-EM without Biggs-Andrews extrapolation (validation)\n"
            "  --precision P     fp32 (default) or mixed: FP16 stored weights, FP32 arithmetic\n"
            "Isotopes:",
            prog);
    for (int i = 0; i < prc_isotope_count(); ++i) fprintf(stream, " %s", prc_isotope_name(i));
    fprintf(stream, "\n");
}

int parse_int(const char* text, const char* name) {  // prc.cu:486-495
    char* end = nullptr;
    errno = 0;
    const long value = strtol(text, &end, 10);
    if (errno || end == text || *end || value < INT_MIN || value > INT_MAX) {
        fprintf(stderr, "Invalid integer for %s: %s\n", name, text);
        exit(EXIT_FAILURE);
    }
    return (int)value;
}

float parse_float(const char* text, const char* name) {  // prc.cu:497-506
    char* end = nullptr;
    errno = 0;
    const float value = strtof(text, &end);
    if (errno || end == text || *end || !std::isfinite(value)) {
        fprintf(stderr, "Invalid finite number for %s: %s\n", name, text);
        exit(EXIT_FAILURE);
    }
    return value;
}

Options parse(int argc, char** argv) {  // prc.cu:508-555
    Options o;
    if (argc < 8) {
        print_usage(argv[0]);
        exit(EXIT_FAILURE);
    }
    o.isotope = argv[1];
    o.nx = parse_int(argv[2], "NX"), o.ny = parse_int(argv[3], "NY"), o.nz = parse_int(argv[4], "NZ");
    o.dx = parse_float(argv[5], "DX"), o.dy = parse_float(argv[6], "DY"), o.dz = parse_float(argv[7], "DZ");
    for (int i = 8; i < argc; ++i) {
        const std::string key = argv[i];
        auto value = [&]() -> const char* {
            if (i + 1 >= argc) {
                fprintf(stderr, "Missing value for %s\n", key.c_str());
                exit(EXIT_FAILURE);
            }
            return argv[++i];
        };
        if (key == "--mode") o.mode = value();
        else if (key == "--input") o.input = value();
        else if (key == "--density") o.density = value();
        else if (key == "--output") o.output = value();
        else if (key == "--rays") o.rays = parse_int(value(), key.c_str());
        else if (key == "--iterations") o.iterations = parse_int(value(), key.c_str());
        else if (key == "--beta") o.beta = parse_float(value(), key.c_str());
        else if (key == "--gamma") o.gamma = parse_float(value(), key.c_str());
        else if (key == "--support-rho") o.support_rho = parse_float(value(), key.c_str());
        else if (key == "--air-cap-mm") o.air_cap_mm = parse_float(value(), key.c_str());
        else if (key == "--save-every") o.save_every = parse_int(value(), key.c_str());
        else if (key == "--precision") o.precision = value();
        else if (key == "--reproject") o.reproject = true;
        else if (key == "--no-cache-near") o.cache_near = false;
        else if (key == "--exact-range") o.exact_range = true;
        else if (key == "--no-accel") o.accelerate = false;
        else {
            fprintf(stderr, "Unknown option %s\n", key.c_str());
            print_usage(argv[0]);
            exit(EXIT_FAILURE);
        }
    }
    if (o.nx <= 0 || o.ny <= 0 || o.nz <= 0 || o.dx <= 0 || o.dy <= 0 || o.dz <= 0 || o.rays < 2 || (o.rays & 1) ||
        o.rays > INT_MAX - 31 || o.iterations < 0 || o.iterations == INT_MAX || o.beta < 0 || o.gamma < 0 ||
        o.support_rho < 0 || o.air_cap_mm <= 0 || o.save_every < 0 ||
        (o.mode != "correct" && o.mode != "forward" && o.mode != "adjoint") ||
        (o.precision != "fp32" && o.precision != "mixed")) {
        fprintf(stderr, "Invalid geometry or solver parameters\n");
        exit(EXIT_FAILURE);
    }
    // Voxel indices are signed 32-bit; check products before multiplying them.
    if (o.nx > INT_MAX / o.ny || o.nx * o.ny > INT_MAX / o.nz) {
        fprintf(stderr, "Volume exceeds the INT_MAX voxel index limit\n");
        exit(EXIT_FAILURE);
    }
    if (o.output.empty())
        o.output = o.mode == "correct" ? "ACTIVITY_PRC_V2.raw" : o.mode == "forward" ? "ACTIVITY_FWD.raw" : "ACTIVITY_ADJ.raw";
    return o;
}

// Insert a suffix before the extension of the file name, never inside a directory name
// (prc.cu:557-562).
std::string with_suffix(const std::string& path, const std::string& suffix) {
    const size_t dot = path.rfind('.');
    const size_t slash = path.find_last_of("/\\");
    const size_t start = slash == std::string::npos ? 0 : slash + 1;
    return dot == std::string::npos || dot <= start ? path + suffix : path.substr(0, dot) + suffix + path.substr(dot);
}

void read_raw(const std::string& path, std::vector<float>* v) {  // prc.cu:339-347
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp || fread(v->data(), sizeof(float), v->size(), fp) != v->size() || fgetc(fp) != EOF || ferror(fp)) {
        fprintf(stderr, "Expected exactly %zu float32 values in %s\n", v->size(), path.c_str());
        exit(EXIT_FAILURE);
    }
    fclose(fp);
}

// Reject non-finite values before opening the file (prc.cu:349-360).
void save_image(const std::string& path, const std::vector<float>& host) {
    for (float value : host)
        if (!std::isfinite(value)) {
            fprintf(stderr, "Non-finite output: %s\n", path.c_str());
            exit(EXIT_FAILURE);
        }
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp || fwrite(host.data(), sizeof(float), host.size(), fp) != host.size() || fclose(fp) != 0) {
        fprintf(stderr, "Cannot write %s\n", path.c_str());
        exit(EXIT_FAILURE);
    }
}

void check(int rc) {
    if (rc != PRC_OK) {
        fprintf(stderr, "prc: %s\n", prc_last_error());
        exit(EXIT_FAILURE);
    }
}

double sum_of(const std::vector<float>& v) {
    double s = 0.0;
    for (float x : v) s += x;
    return s;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 2 && (!strcmp(argv[1], "--help") || !strcmp(argv[1], "-h"))) {
        print_usage(argv[0], stdout);
        return EXIT_SUCCESS;
    }
    const Options o = parse(argc, argv);
    const int iso = prc_isotope_find(o.isotope.c_str());
    if (iso < 0) {
        print_usage(argv[0]);
        return EXIT_FAILURE;
    }
    const float cutoff = prc_range_cutoff(iso, 0.9995f);  // prc.cu:587
    const size_t n = (size_t)o.nx * o.ny * o.nz;

    const double t0 = now_s();
    std::vector<float> rho(n), img(n);
    read_raw(o.density, &rho);
    read_raw(o.input, &img);
    size_t n_active = 0;
    double input_sum = 0.0, support_sum = 0.0;
    for (size_t i = 0; i < n; ++i) {  // prc.cu:598-613
        if (!std::isfinite(rho[i]) || !std::isfinite(img[i])) {
            fprintf(stderr, "Non-finite density or activity at voxel %zu\n", i);
            return EXIT_FAILURE;
        }
        rho[i] = std::max(rho[i], kMinDensity);
        img[i] = std::max(img[i], 0.0f);
        input_sum += img[i];
        if (rho[i] >= o.support_rho) {
            ++n_active;
            support_sum += img[i];
        } else {
            img[i] = 0.0f;
        }
    }
    printf("isotope %s, grid %dx%dx%d, voxel %.3fx%.3fx%.3f mm, cutoff %.3f mm (water), air cap %.1f mm\n",
           o.isotope.c_str(), o.nx, o.ny, o.nz, o.dx, o.dy, o.dz, cutoff, o.air_cap_mm);
    printf("tissue voxels %zu (%.1f%%), input counts in tissue %.6g of %.6g (%.2f%%)\n", n_active,
           100.0 * n_active / n, support_sum, input_sum, input_sum > 0 ? 100.0 * support_sum / input_sum : 0.0);
    if (support_sum == 0.0) {
        fprintf(stderr, "No activity in voxels with density >= %g: nothing to do\n", o.support_rho);
        return EXIT_FAILURE;
    }

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
    printf("system matrix: %lld stored weights (%.0f MiB, %s)\n", (long long)st.weight_entries,
           st.weight_bytes / 1048576.0, o.precision.c_str());
    printf("normalisation (%d near-block + %d far-field rays): %.2f s\n", PRC_NEAR_RAYS, o.rays, now_s() - t);

    std::vector<float>& out = img;  // reuse the input buffer for outputs
    if (o.mode == "forward" || o.mode == "adjoint") {
        t = now_s();
        std::vector<float> result(n, 0.0f);
        if (o.mode == "forward") check(prc_projector_forward_host(A, img.data(), result.data()));
        else check(prc_projector_adjoint_host(A, img.data(), result.data()));
        printf("%s: %.3f s\n", o.mode.c_str(), now_s() - t);
        save_image(o.output, result);
        printf("sum in %.6g, sum out %.6g, total %.2f s\n", support_sum, sum_of(result), now_s() - t0);
        prc_projector_destroy(A);
        return EXIT_SUCCESS;
    }

    printf("MAP-EM: %d iterations, beta %.4g, gamma %.3g\n", o.iterations, o.beta, o.gamma);
    prc_solve_params p{o.beta, o.gamma, o.accelerate ? 1 : 0};
    prc_solver* S = nullptr;
    check(prc_solver_create(A, img.data(), &p, &S));
    t = now_s();
    for (int it = 1; it <= o.iterations;) {
        // Run up to the next iteration that prints or saves (prc.cu:701-709).
        int next = it;
        while (next < o.iterations && !(next == 1 || next % 10 == 0) &&
               !(o.save_every > 0 && next % o.save_every == 0))
            ++next;
        check(prc_solver_iterate(S, next - it + 1));
        if (next == 1 || next % 10 == 0 || next == o.iterations) {
            check(prc_solver_sync(S));
            printf("iteration %d/%d: %.2f s\n", next, o.iterations, now_s() - t);
        }
        if (o.save_every > 0 && next % o.save_every == 0 && next < o.iterations) {
            char tag[32];
            snprintf(tag, sizeof(tag), "_it%03d", next);
            check(prc_solver_estimate(S, out.data()));
            save_image(with_suffix(o.output, tag), out);
        }
        it = next + 1;
    }
    check(prc_solver_sync(S));
    const double t_iter = now_s() - t;
    check(prc_solver_estimate(S, out.data()));
    save_image(o.output, out);
    const double corrected_sum = sum_of(out);
    if (o.reproject) {
        check(prc_solver_reproject(S, out.data()));
        save_image(with_suffix(o.output, "_REPROJ"), out);
    }
    printf("sum blurred (tissue) %.6g, sum corrected %.6g\n", support_sum, corrected_sum);
    printf("iterations %.2f s (%.2f s each), total %.2f s\n", t_iter, o.iterations ? t_iter / o.iterations : 0.0,
           now_s() - t0);
    prc_solver_destroy(S);
    prc_projector_destroy(A);
    return EXIT_SUCCESS;
}
