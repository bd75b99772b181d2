// Driver for the C++ API tests (run by tests/test_cpp_api.py).
//
//   test_api cpu <golden>                   host-only checks (VCF input, result, errors)
//   test_api fixture <golden> <out>         Imputer on test_reference/test_target.vcf
//   test_api run <in> <out> <L> <ne> <per_marker> <argmax> <batch> <seed> <use_pbwt>
//            inputs: in/meta.txt (H M B), in/cols.bin, in/pos.bin, in/cm.bin, in/codes.bin,
//            in/palette.bin; outputs: out/{prefix,div,states,hap,ap,loglik}.bin
// Mirrors the reference's executable tests (test/test_vcf_io.cpp: load the
// fixtures, build PBWT, select states, write VCF).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/swiftimpute/imputer.hpp"

using namespace swiftimpute;

static int failures = 0;
#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);              \
            ++failures;                                                              \
        }                                                                            \
    } while (0)

template <class T>
static std::vector<T> read_bin(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot read " + path);
    const size_t n = size_t(f.tellg());
    f.seekg(0);
    std::vector<T> v(n / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
    return v;
}

template <class T>
static void write_bin(const std::string& path, const T* p, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), std::streamsize(n * sizeof(T)));
}

static int cpu_checks(const std::string& golden) {
    auto ref = ReferencePanel::load_vcf(golden + "/test_reference.vcf");
    CHECK(ref->num_markers() == 10);
    CHECK(ref->num_samples() == 4);
    CHECK(ref->num_haplotypes() == 8);
    CHECK(ref->get_allele(0, 3) == 1 && ref->get_allele(0, 2) == 0);
    CHECK(ref->haplotypes()[1 * 8 + 4] == 1);
    auto tgt = TargetData::load_vcf(golden + "/test_target.vcf");
    CHECK(tgt->num_markers() == 10);
    CHECK(tgt->num_samples() == 2);
    auto g = tgt->get_likelihood(0, 0);  // 0/1
    CHECK(g.ll_00 == -999.0f && g.ll_01 == 0.0f && g.ll_11 == -999.0f);
    g = tgt->get_likelihood(0, 1);  // ./.
    CHECK(g.ll_00 == 0.0f && g.ll_01 == 0.0f && g.ll_11 == 0.0f);
    g = tgt->get_likelihood(1, 0);  // 1/1
    CHECK(g.ll_11 == 0.0f && g.ll_00 == -999.0f);
    g = tgt->get_likelihood(1, 6);  // 1/0 -> het
    CHECK(g.ll_01 == 0.0f);
    CHECK(tgt->palette().size() == 4);
    CHECK(tgt->code_pitch() == 16);
    auto rg = ReferencePanel::load_vcf(golden + "/test_reference.vcf", "chr1:2000-4000");
    CHECK(rg->num_markers() == 3);
    bool threw = false;
    try {
        ReferencePanel::load_vcf(golden + "/nope.vcf.gz");
    } catch (const ImputationError&) {
        threw = true;
    }
    CHECK(threw);
    // result accessors, INFO and the writer (imputer.cpp:333-511)
    ImputationResult r(2, 10);
    std::vector<allele_t> h0(10, 0), h1(10, 1);
    r.set_haplotype(0, 0, h0.data());
    r.set_haplotype(0, 1, h1.data());
    r.set_haplotype(1, 0, h1.data());
    r.set_haplotype(1, 1, h1.data());
    CHECK(r.get_dosage(0, 3) == 1.0f && r.get_dosage(1, 3) == 2.0f);
    // without posteriors INFO comes from the sampled dosages, as the reference:
    // dosages 1 and 2: mean 1.5, p 0.75, var 0.25, 2p(1-p) 0.375 -> INFO 1/3
    CHECK(!r.has_posteriors());
    r.compute_info_scores();
    CHECK(std::fabs(r.get_info_score(0) - (1.0f - 0.25f / 0.375f)) < 1e-6);
    ImputationConfig cfg;
    CHECK(cfg.batch_size == 100 && cfg.window_size == 10000 && cfg.hmm_params.num_states == 8 &&
          cfg.hmm_params.ne == 10000 && cfg.hmm_params.rho_rate == 4.0f);
    const std::string out = "/tmp/si_test_output.vcf";
    auto first_line = [&](const std::string& path, int* n) {
        std::ifstream f(path);
        std::string line, first;
        int variants = 0;
        while (std::getline(f, line))
            if (!line.empty() && line[0] != '#') {
                if (!variants) first = line;
                ++variants;
            }
        *n = variants;
        return first;
    };
    auto header_has = [&](const std::string& path, const std::string& what) {
        std::ifstream f(path);
        std::string line;
        while (std::getline(f, line))
            if (line.rfind("##", 0) == 0 && line.find(what) != std::string::npos) return true;
        return false;
    };
    int variants = 0;
    // no posteriors: GT only, and only GT declared
    r.write_vcf(out, *tgt, *ref, cfg);
    CHECK(first_line(out, &variants) == "chr1\t1000\trs1\tA\tG\t.\tPASS\tIMP;INFO=0.3333\tGT\t0|1\t1|1");
    CHECK(variants == 10 && !header_has(out, "ID=DS") && header_has(out, "ID=INFO"));
    // posterior allele probabilities at every (sample, marker)
    for (sample_t s = 0; s < 2; ++s)
        for (marker_t m = 0; m < 10; ++m) {
            const prob_t ap[2] = {s == 0 ? 0.25f : 0.5f, 0.5f};
            r.set_posterior(s, m, ap);
        }
    CHECK(r.has_posteriors());
    prob_t gp[3];
    r.get_probabilities(0, 2, gp);
    CHECK(std::fabs(gp[0] - 0.375f) < 1e-7 && std::fabs(gp[1] - 0.5f) < 1e-7 && std::fabs(gp[2] - 0.125f) < 1e-7);
    CHECK(std::fabs(r.get_posterior_dosage(0, 2) - 0.75f) < 1e-7);
    // INFO now from the posterior dosages 0.75 and 1.0: mean 0.875, p 0.4375,
    // var 0.015625, 2p(1-p) 0.4921875
    r.compute_info_scores();
    CHECK(std::fabs(r.get_info_score(0) - float(1.0 - 0.015625 / 0.4921875)) < 1e-6);
    r.write_vcf(out, *tgt, *ref, cfg);
    CHECK(first_line(out, &variants) ==
          "chr1\t1000\trs1\tA\tG\t.\tPASS\tIMP;INFO=0.9683\tGT:DS:GP:AP"
          "\t0|1:0.7500:0.3750,0.5000,0.1250:0.2500,0.5000\t1|1:1.0000:0.2500,0.5000,0.2500:0.5000,0.5000");
    CHECK(variants == 10 && header_has(out, "ID=DS") && header_has(out, "ID=GP") && header_has(out, "ID=AP"));
    // the output flags select the fields; with all off the line is the reference's
    ImputationConfig c3;
    c3.output_dosages = false;
    r.write_vcf(out, *tgt, *ref, c3);
    CHECK(first_line(out, &variants).find("\tGT:GP:AP\t0|1:0.3750,0.5000,0.1250:0.2500,0.5000\t") !=
          std::string::npos);
    c3.output_probabilities = false;
    c3.output_info_score = false;
    r.write_vcf(out, *tgt, *ref, c3);
    CHECK(first_line(out, &variants) == "chr1\t1000\trs1\tA\tG\t.\tPASS\tIMP\tGT\t0|1\t1|1");
    CHECK(!header_has(out, "ID=DS") && !header_has(out, "ID=GP") && !header_has(out, "ID=INFO"));
    // a missing panel allele reaches GT as "."
    std::vector<allele_t> hm(10, 255);
    r.set_haplotype(1, 0, hm.data());
    r.write_vcf(out, *tgt, *ref, c3);
    CHECK(first_line(out, &variants) == "chr1\t1000\trs1\tA\tG\t.\tPASS\tIMP\tGT\t0|1\t.|1");
    CHECK(Imputer::estimate_memory_requirement(*ref, 2, cfg) > 0);
    bool bad = false;
    try {
        std::vector<allele_t> a = {0, 2};
        std::vector<Marker> mk(1);
        ReferencePanel::from_alleles(mk, {}, 2, a.data());
    } catch (const ImputationError&) {
        bad = true;
    }
    CHECK(bad);
    std::printf("%s cpu checks (%d failures)\n", failures ? "FAIL" : "PASS", failures);
    return failures ? 1 : 0;
}

static int fixture(const std::string& golden, const std::string& out) {
    auto ref = ReferencePanel::load_vcf(golden + "/test_reference.vcf");
    auto tgt = TargetData::load_vcf(golden + "/test_target.vcf");
    ImputationConfig cfg;
    cfg.hmm_params.num_states = 4;
    Imputer imp(*ref, cfg);
    CHECK(!imp.has_index());
    imp.build_index();
    CHECK(imp.has_index());
    auto prefix = imp.pbwt_prefix();
    CHECK(prefix.size() == 80);
    const uint32_t kat0[8] = {0, 1, 2, 5, 3, 4, 6, 7}, kat1[8] = {0, 1, 2, 3, 5, 4, 6, 7};
    for (int i = 0; i < 8; ++i) CHECK(prefix[i] == kat0[i] && prefix[8 + i] == kat1[i]);
    auto div = imp.pbwt_divergence();
    for (uint32_t m = 0; m < 10; ++m)
        for (int i = 0; i < 8; ++i) CHECK(div[m * 8 + i] == m);
    auto st = imp.selected_states();
    CHECK(st[0] == 0 && st[1] == 1 && st[2] == 2 && st[3] == 5);  // select_states(0, ..., 4)
    auto res = imp.impute(*tgt);
    for (sample_t s = 0; s < 2; ++s)
        for (marker_t m = 0; m < 10; ++m) {
            CHECK(res->haplotype0(s)[m] <= 1 && res->haplotype1(s)[m] <= 1);
            const float d = res->get_posterior_dosage(s, m);
            CHECK(d >= 0.0f && d <= 2.0f);
        }
    res->write_vcf(out + "/fixture_out.vcf", *tgt, *ref, cfg);
    // marker-count mismatch is an ImputationError (imputer.cpp:730-741)
    bool threw = false;
    try {
        auto short_tgt = TargetData::load_vcf(golden + "/test_target.vcf", "chr1:1000-5000");
        imp.impute(*short_tgt);
    } catch (const ImputationError&) {
        threw = true;
    }
    CHECK(threw);
    // without build_index(): simple selection pairs (2k, 2k+1)
    ImputationConfig c2;
    c2.hmm_params.num_states = 4;
    Imputer imp2(*ref, c2);
    auto r2 = imp2.impute(*tgt);
    auto st2 = imp2.selected_states();
    CHECK(st2[0] == 0 && st2[1] == 2 && st2[2] == 4 && st2[3] == 6);
    CHECK(r2->num_samples() == 2);
    write_bin(out + "/fixture_hap.bin", res->haplotype0(0), size_t(2) * 2 * 10);
    // --region: the panel and targets of a region, the index swept from the
    // identity order at its first marker (main.cpp:103, imputer.cpp:40-44)
    auto rref = ReferencePanel::load_vcf(golden + "/test_reference.vcf", "chr1:4000-8000");
    auto rtgt = TargetData::load_vcf(golden + "/test_target.vcf", "chr1:4000-8000");
    CHECK(rref->num_markers() == 5 && rtgt->num_markers() == 5);
    Imputer rimp(*rref, cfg);
    rimp.build_index();
    auto rp = rimp.pbwt_prefix();
    auto rs = rimp.selected_states();
    write_bin(out + "/region_prefix.bin", rp.data(), rp.size());
    write_bin(out + "/region_states.bin", rs.data(), rs.size());
    auto rres = rimp.impute(*rtgt);
    CHECK(rres->num_samples() == 2 && rres->num_markers() == 5);
    std::printf("%s fixture checks (%d failures)\n", failures ? "FAIL" : "PASS", failures);
    return failures ? 1 : 0;
}

static int run(int argc, char** argv) {
    const std::string in = argv[2], out = argv[3];
    const uint32_t L = std::stoul(argv[4]), ne = std::stoul(argv[5]);
    const bool per_marker = std::stoi(argv[6]) != 0, argmax = std::stoi(argv[7]) != 0;
    const uint32_t batch = std::stoul(argv[8]);
    const uint64_t seed = std::stoull(argv[9]);
    const bool use_pbwt = argc > 10 ? std::stoi(argv[10]) != 0 : true;
    const bool gpu_compat = argc > 11 ? std::stoi(argv[11]) != 0 : false;
    const uint32_t window = argc > 12 ? std::stoul(argv[12]) : 10000u;
    uint32_t H, M, B;
    {
        std::ifstream meta(in + "/meta.txt");
        meta >> H >> M >> B;
    }
    auto cols = read_bin<uint64_t>(in + "/cols.bin");
    auto pos = read_bin<uint64_t>(in + "/pos.bin");
    auto cm = read_bin<double>(in + "/cm.bin");
    auto codes = read_bin<uint8_t>(in + "/codes.bin");  // [B][M]
    auto palf = read_bin<float>(in + "/palette.bin");
    std::vector<Marker> markers(M);
    for (uint32_t m = 0; m < M; ++m) {
        markers[m].chrom = "chr1";
        markers[m].pos = pos[m];
        markers[m].cM = cm[m];
    }
    std::vector<GenotypeLikelihoods> pal;
    for (size_t p = 0; p + 2 < palf.size(); p += 3) pal.emplace_back(palf[p], palf[p + 1], palf[p + 2]);
    auto ref = ReferencePanel::from_bitpacked(markers, H, cols.data());
    auto tgt = TargetData::from_codes(markers, B, pal, codes.data());
    ImputationConfig cfg;
    cfg.hmm_params.num_states = L;
    cfg.hmm_params.ne = ne;
    cfg.hmm_params.use_pbwt_selection = use_pbwt;
    cfg.selection = gpu_compat ? SelectionMode::GpuCompat : SelectionMode::CpuSort;
    cfg.window_size = window;
    cfg.transition_mode = per_marker ? TransitionMode::PerMarker : TransitionMode::FirstInterval;
    cfg.sampling = argmax ? SamplingMode::Argmax : SamplingMode::Philox;
    cfg.batch_size = batch;
    cfg.seed = seed;
    Imputer imp(*ref, cfg);
    imp.build_index();
    uint32_t calls = 0, last = 0;
    auto res = imp.impute_with_progress(*tgt, [&](uint32_t done, uint32_t total) {
        ++calls;
        last = done;
        CHECK(total == B);
    });
    CHECK(calls == (B + batch - 1) / batch && last == B);
    // a second call reuses the device buffers: the reported device memory
    // stays the same and the results are identical
    const size_t mem1 = imp.device_memory_usage();
    auto res2 = imp.impute(*tgt);
    CHECK(imp.device_memory_usage() == mem1);
    CHECK(std::memcmp(res2->haplotype0(0), res->haplotype0(0), size_t(B) * 2 * M) == 0);
    // a rebuild of the index keeps the accounting (arrays kept or reused)
    imp.build_index();
    CHECK(imp.device_memory_usage() == mem1);
    auto prefix = imp.pbwt_prefix();
    auto div = imp.pbwt_divergence();
    auto st = imp.selected_states();
    write_bin(out + "/prefix.bin", prefix.data(), prefix.size());
    write_bin(out + "/div.bin", div.data(), div.size());
    write_bin(out + "/states.bin", st.data(), st.size());
    write_bin(out + "/hap.bin", res->haplotype0(0), size_t(B) * 2 * M);
    write_bin(out + "/ap.bin", res->mutable_posteriors(), size_t(B) * M * 2);
    std::vector<double> ll(B);
    for (uint32_t s = 0; s < B; ++s) ll[s] = res->log_likelihood(s);
    write_bin(out + "/loglik.bin", ll.data(), ll.size());
    std::printf("%s run (%d failures), device memory %zu bytes\n", failures ? "FAIL" : "PASS", failures,
                imp.device_memory_usage());
    return failures ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: test_api cpu|fixture|run ...\n");
        return 2;
    }
    const std::string mode = argv[1];
    try {
        if (mode == "cpu") return cpu_checks(argv[2]);
        if (mode == "fixture" && argc >= 4) return fixture(argv[2], argv[3]);
        if (mode == "run" && argc >= 10) return run(argc, argv);
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
