// Drives the C++ drop-in classes the way the reference harness does
// (src/OCLTest.cpp:48-130): level-shifted planar components, OCLEncoder::run,
// finish, mapDWTOut.  Reads a raw interleaved RGB uint8 image, writes the
// mapped outputs for the Python test to compare with the CPU oracle.
//   dropin_main <in.rgb> <w> <h> <outdir>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "roger/OCLDecoder.h"
#include "roger/OCLDeviceManager.h"
#include "roger/OCLEncoder.h"

static bool write_file(const std::string &path, const void *p, size_t n)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = fwrite(p, 1, n, f) == n;
    fclose(f);
    return ok;
}

int main(int argc, char **argv)
{
    if (argc != 5) { fprintf(stderr, "usage: %s in.rgb w h outdir\n", argv[0]); return 2; }
    const size_t w = atoi(argv[2]), h = atoi(argv[3]), n = w * h;
    std::string out = argv[4];
    std::vector<unsigned char> rgb(n * 3);
    FILE *f = fopen(argv[1], "rb");
    if (!f || fread(rgb.data(), 1, rgb.size(), f) != rgb.size()) { fprintf(stderr, "read failed\n"); return 2; }
    fclose(f);

    OCLDeviceManager dm;
    if (dm.init() != DeviceSuccess) return 3;
    ocl_args_d_t *ocl = dm.getInfo();

    // 4 channels, A duplicating B, level shifted by the caller (src/OCLTest.cpp:81-90)
    std::vector<std::vector<float>> fch(4, std::vector<float>(n));
    std::vector<std::vector<short>> sch(4, std::vector<short>(n));
    for (int c = 0; c < 4; ++c) {
        int src = c == 3 ? 2 : c;
        for (size_t i = 0; i < n; ++i) {
            fch[c][i] = (float)((int)rgb[i * 3 + src] - 128);
            sch[c][i] = (short)((int)rgb[i * 3 + src] - 128);
        }
    }
    std::vector<float *> fcomp;
    std::vector<short *> scomp;
    for (int c = 0; c < 4; ++c) { fcomp.push_back(fch[c].data()); scomp.push_back(sch[c].data()); }
    const size_t levels = 5, precision = 8;

    // lossy, quantised (the harness's OCLTest<float, short>(true, false))
    {
        OCLEncoder<float> enc(ocl, true, false);
        enc.run(fcomp, w, h, levels, precision);
        enc.run(fcomp, w, h, levels, precision);   // second run reuses resident buffers
        enc.finish();
        short *res = nullptr;
        if (enc.mapDWTOut((void **)&res) != DeviceSuccess || !res) return 4;
        if (!write_file(out + "/lossy_q.i16", res, n * 4 * sizeof(short))) return 5;
        enc.unmapDWTOut(res);
        std::vector<rg_cb_desc> d; std::vector<rg_cb_result> r; std::vector<long> o; std::vector<uint8_t> data;
        if (enc.getCodeblocks(d, r, o, data) != DeviceSuccess) return 6;
        if (!write_file(out + "/lossy_cb.res", r.data(), r.size() * sizeof(rg_cb_result))) return 5;
        if (!write_file(out + "/lossy_cb.bin", data.data(), data.size())) return 5;
        // per-channel int16 planes of the planar stage, read through mapBuffer
        OCLMemoryManager<float> *mm = enc.getMemoryManager();
        for (int c = 0; c < 4; ++c) {
            tDeviceMem plane = mm->getDWTOutByChannel(c);
            short *hp = nullptr;
            if (!plane || mm->mapBuffer(plane, (void **)&hp) != DeviceSuccess || !hp) return 7;
            if (!write_file(out + "/planar_c" + std::to_string(c) + ".i16", hp, n * sizeof(short))) return 5;
            if (mm->unmapMemory(plane, hp) != DeviceSuccess) return 7;
        }
        rg_quiet() = true;
        if (mm->getDWTOutByChannel(4)) return 7;
        // the dwtIn pyramid: inputs of levels 0 .. levels-1, nothing beyond
        for (size_t l = 0; l < levels; ++l)
            if (!mm->getDwtIn(l)) return 13;
        if (mm->getDwtIn(levels)) return 13;
        rg_quiet() = false;
        // LL after one level (input of level 2), component 0, float
        {
            const size_t w1 = (w + 1) / 2, h1 = (h + 1) / 2;
            std::vector<float> ll(w1 * h1);
            if (rg_memcpy_d2h(ocl->context, ll.data(), mm->getDwtIn(1), ll.size() * sizeof(float)) != RG_SUCCESS)
                return 13;
            if (!write_file(out + "/dwtin1_c0.f32", ll.data(), ll.size() * sizeof(float))) return 5;
        }
    }
    // lossy float DWT out, then the decoder reconstructs the samples
    {
        OCLEncoder<float> enc(ocl, true, true);
        enc.run(fcomp, w, h, levels, precision);
        enc.finish();
        float *coef = nullptr;
        if (enc.mapDWTOut((void **)&coef) != DeviceSuccess) return 8;
        std::vector<std::vector<float>> planes(4, std::vector<float>(n));
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 4; ++c) planes[c][i] = coef[i * 4 + c];
        enc.unmapDWTOut(coef);
        if (!write_file(out + "/lossy_coef.f32", planes[0].data(), n * 4)) return 5;
        OCLDecoder<float> dec(ocl, true);
        std::vector<float *> pp;
        for (int c = 0; c < 4; ++c) pp.push_back(planes[c].data());
        dec.run(pp, w, h, levels, precision);
        dec.finish();
        float *rec = nullptr;
        if (dec.mapDWTOut((void **)&rec) != DeviceSuccess) return 9;
        if (!write_file(out + "/lossy_rec.f32", rec, n * 4 * sizeof(float))) return 5;
        dec.unmapDWTOut(rec);
    }
    // lossless 5/3 round trip must be exact
    {
        OCLEncoder<short> enc(ocl, false, false);
        enc.run(scomp, w, h, levels, precision);
        enc.finish();
        short *coef = nullptr;
        if (enc.mapDWTOut((void **)&coef) != DeviceSuccess) return 10;
        std::vector<std::vector<short>> planes(4, std::vector<short>(n));
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 4; ++c) planes[c][i] = coef[i * 4 + c];
        enc.unmapDWTOut(coef);
        if (!write_file(out + "/lossless_coef.i16", planes[0].data(), n * 2)) return 5;
        OCLDecoder<short> dec(ocl, false);
        std::vector<short *> pp;
        for (int c = 0; c < 4; ++c) pp.push_back(planes[c].data());
        dec.run(pp, w, h, levels, precision);
        dec.finish();
        short *rec = nullptr;
        if (dec.mapDWTOut((void **)&rec) != DeviceSuccess) return 11;
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 4; ++c)
                if (rec[i * 4 + c] != scomp[c][i]) { fprintf(stderr, "lossless mismatch at %zu/%d\n", i, c); return 12; }
        dec.unmapDWTOut(rec);
    }
    printf("dropin ok %zux%zu\n", w, h);
    return 0;
}
