// Host replay of the array indices one resident migrate launch forms
// (migrate_tile in kirch_datuming_amd/csrc/kmig_migrate.hip), through the same
// index functions (kirch_datuming_amd/csrc/kmig_index.h), checked against the
// sizes the library allocates for them (kirch_datuming_amd/csrc/kmig_api.cpp).
// It walks the launch's grid as the kernel does (tiles, waves, shot groups,
// receiver chunks or columns, shots, receivers of a chunk) and visits every
// shot of a chunk's range, as if no group were culled, so it covers every
// index any cull decision can leave; and the survivor-mask words the launch's
// k_cull_mask writes and the launch reads.  Built with g++ by
// tests/test_bounds_replay.py.
//
// usage: bounds_replay nz nx TZ TX R ns nh nt sb se nsplit cols unclamped tile_step
//   cols       0: fixed receivers; 1: receivers moving with the shot on a
//              lattice (receiver column c = batch shot + receiver index)
//   unclamped  1: read padding receivers' tap rows without the clamp to the
//              shot's last trace (the addressing that once faulted)
//   tile_step  visit every tile_step-th tile (1: all), plus the last tile
// Prints one line per array (largest index + 1 and the allocated size) and
// "violations N".
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kmig_index.h"
#include "kmig_internal.h"

using namespace kmig;

namespace {

struct Arr {
    const char *name;
    size_t size;
    size_t hi = 0;  // largest index + 1 seen
    long bad = 0;
    void at(size_t k)
    {
        hi = std::max(hi, k + 1);
        if (k >= size) bad++;
    }
};

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 15) {
        fprintf(stderr, "usage: %s nz nx TZ TX R ns nh nt sb se nsplit cols unclamped tile_step\n",
                argv[0]);
        return 2;
    }
    int v[14];
    for (int k = 0; k < 14; k++) v[k] = atoi(argv[k + 1]);
    const int nz = v[0], nx = v[1], TZ = v[2], TX = v[3], R = v[4], ns = v[5], nh = v[6],
              nt = v[7], sb = v[8], se = v[9], nsplit = v[10], cols = v[11], unclamped = v[12],
              tile_step = std::max(1, v[13]);
    const size_t nzx = (size_t)nz * nx;
    const int nsb = se - sb;
    // receiver columns on a lattice: trace (js, ih) sits in column js + ih
    const int ncols = cols ? nsb + nh - 1 : 0;
    const int nrec = cols ? ncols : nh;
    const int nrslots = nrec;  // one tuple slot per receiver index or column
    std::vector<ColTrace> ctab;
    std::vector<int> cfirst, clast;
    if (cols) {
        ctab.assign((size_t)nsb * ncols, ColTrace{0, -1.0f});
        ctab.resize((size_t)nsb * ncols + RES_WAP_PAD, ColTrace{0, 0.0f});
        cfirst.assign(ncols, 1 << 30);
        clast.assign(ncols, -1);
        for (int js = 0; js < nsb; js++)
            for (int ih = 0; ih < nh; ih++) {
                const int c = js + ih;
                ctab[(size_t)js * ncols + c] = ColTrace{ih, 1.0f};
                cfirst[c] = std::min(cfirst[c], js);
                clast[c] = std::max(clast[c], js);
            }
    }
    const int ncz = (nz + BND_CZ - 1) / BND_CZ, ncx = (nx + BND_CX - 1) / BND_CX;
    const size_t ncell = (size_t)ncz * ncx;
    const int nchunk = (nrec + R - 1) / R, nwords = (nsb + 63) / 64;
    // sizes as kmig_api.cpp allocates them (masks: for chunks of 4 receivers)
    Arr pairs{"pairs", (size_t)ns * nh * nt}, wap{"wap", (size_t)ns * nh + RES_WAP_PAD},
        ctabA{"ctab", ctab.size()}, sbnd{"sbnd", ncell * nsb},
        mask{"mask", ncell * (size_t)((nrec + 3) / 4) * (size_t)((std::max(1, nsb) + 63) / 64)},
        rbnd{"rbnd", ncell * nrslots}, stup{"stup", (size_t)nsb * nzx},
        rtup{"rtup", (size_t)nrslots * nzx}, rslot{"rslot", (size_t)nrec + RSLOT_PAD},
        cfA{"cfirst", (size_t)ncols}, clA{"clast", (size_t)ncols};
    // k_cull_mask: block (cell, word), lane = shot, a loop over the chunks
    for (size_t c = 0; c < ncell; c++)
        for (int w = 0; w < nwords; w++) {
            for (int lane = 0; lane < 64; lane++)
                if (w * 64 + lane < nsb) sbnd.at(idx::bnd((int)c, (size_t)nsb, w * 64 + lane));
            for (int k = 0; k < nchunk; k++) {
                for (int lane = 0; lane < R && k * R + lane < nrec; lane++) {
                    const int ih = k * R + lane;
                    rslot.at(ih);
                    rbnd.at(idx::bnd((int)c, (size_t)nrslots, ih));  // slot = rslot[ih]
                    if (cols) {
                        cfA.at(ih);
                        clA.at(ih);
                    }
                }
                mask.at(idx::mask_word((int)c, nchunk, k, nwords, w));
            }
        }
    const int ntz = (nz + TZ - 1) / TZ, ntx = (nx + TX - 1) / TX, ntiles = ntz * ntx;
    const int WX = 64 / TZ, NCZ = (TZ + BND_CZ - 1) / BND_CZ, NCX = (WX + BND_CX - 1) / BND_CX;
    const int nwave = TZ * TX / 64;
    std::vector<int> cell(NCZ * NCX);
    for (int tile = 0; tile < ntiles; tile = (tile + tile_step < ntiles || tile == ntiles - 1)
                                                 ? tile + tile_step
                                                 : ntiles - 1) {
        const int bz = tile % ntz, bx = tile / ntz;
        for (int wv = 0; wv < nwave; wv++) {
            // the wave's points: largest flat index of a valid lane (invalid
            // lanes read index 0)
            size_t imax = 0;
            for (int l = wv * 64; l < wv * 64 + 64; l++) {
                const int iz = bz * TZ + l % TZ, ix = bx * TX + l / TZ;
                if (iz < nz && ix < nx) imax = std::max(imax, (size_t)ix * nz + iz);
            }
            const int cz0 = (bz * TZ) / BND_CZ, cx0 = (bx * TX + wv * WX) / BND_CX;
            for (int jx = 0; jx < NCX; jx++)
                for (int jz = 0; jz < NCZ; jz++)
                    cell[jx * NCZ + jz] = idx::cell(cx0, jx, cz0, jz, ncx, ncz);
            for (int sg = 0; sg < nsplit; sg++) {
                const int jb = (int)((long)sg * nsb / nsplit), je = (int)((long)(sg + 1) * nsb / nsplit);
                for (int r0 = 0; r0 < nrec; r0 += R) {
                    // the chunk's survivor words over the shot group
                    if (je > jb)
                        for (int c : cell)
                            for (int w = jb >> 6; w <= (je - 1) >> 6; w++)
                                mask.at(idx::mask_word(c, nchunk, r0 / R, nwords, w));
                    int cjb = jb, cje = je;
                    if (cols) {
                        int f = 1 << 30, l = -1;
                        for (int r = 0; r < R && r0 + r < nrec; r++) {
                            cfA.at(r0 + r);
                            clA.at(r0 + r);
                            f = std::min(f, cfirst[r0 + r]);
                            l = std::max(l, clast[r0 + r]);
                        }
                        cjb = std::max(jb, f);
                        cje = std::min(je, l + 1);
                        if (cjb >= cje) continue;
                    }
                    for (int r = 0; r < R; r++) {
                        // every slot of the chunk is read (-1 past the last
                        // receiver, whose tuple load then reads slot 0)
                        const int ih = r0 + r;
                        rslot.at(ih);
                        const int slot = ih < nrec ? ih : 0;  // max(rslot[ih], 0)
                        rtup.at((size_t)slot * nzx + imax);
                    }
                    for (int j = cjb; j < cje; j++) {
                        stup.at((size_t)j * nzx + imax);
                        const int is = sb + j;
                        for (int g = 0; g < R; g++) {
                            size_t row;
                            if (cols) {
                                const size_t k = idx::ctab(0, j, ncols, r0, g);
                                ctabA.at(k);
                                const int ih = k < ctab.size() ? ctab[k].ih : 0;
                                row = idx::col_row(is, nh, ih);
                            } else {
                                wap.at(idx::wap(is, nh, r0, g));
                                row = unclamped ? (size_t)is * nh + r0 + g
                                                : idx::tap_row(is, nh, r0, g);
                            }
                            // taps at pair offsets 0 .. nt - 1 (the last pair
                            // is the sentinel rejected updates read)
                            pairs.at(row * nt + (nt - 1));
                        }
                    }
                }
            }
        }
        if (tile == ntiles - 1) break;
    }
    long bad = 0;
    for (Arr *a : {&pairs, &wap, &ctabA, &sbnd, &mask, &rbnd, &stup, &rtup, &rslot, &cfA, &clA}) {
        printf("%s %zu %zu\n", a->name, a->hi, a->size);
        bad += a->bad;
    }
    printf("violations %ld\n", bad);
    return 0;
}
