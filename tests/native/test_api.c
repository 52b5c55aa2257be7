/* The library API (include/gail_amd.h) driven the way an application drives it, compiled as
 * C: pull (send / receive), callback (on_packet), blocking encodes from several threads,
 * finish, cancel, the end of the stream, and the failure paths (device lost, no device
 * memory, a plugin the loader must refuse). It follows the cases of the reference's
 * tests/test_api.c:106-350 (test_pull, test_callback, test_threads, test_cancel).
 *
 *   test_api <case> [arguments]
 *
 * With the real HIP plugin next to libgail.so the cases run on the GPU; with a stand-in
 * plugin (tests/native/stub_plugin.cpp) they run on the CPU. `pull` can write each frame's
 * input and codestream to a directory, so that a test can recompute the codestream from the
 * same pixels on another path. Exit status 0 when every check held; each failed check prints
 * a FAIL line. A watchdog ends the process (exit 3) if a call does not return. */
#define _GNU_SOURCE 1
#include <pthread.h>
#include <stddef.h>
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "gail_amd.h"

static int failures = 0;
#define CHECK(c, ...)                                     \
    do {                                                  \
        if (!(c)) {                                       \
            printf("  FAIL %s:%d: ", __FILE__, __LINE__); \
            printf(__VA_ARGS__);                          \
            printf("\n");                                 \
            fflush(stdout);                               \
            ++failures;                                   \
        }                                                 \
    } while (0)

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

static void sleep_ms(int ms) {
    struct timespec ts = {ms / 1000, (ms % 1000) * 1000000L};
    nanosleep(&ts, NULL);
}

static void on_alarm(int sig) {
    (void)sig;
    static const char msg[] = "  FAIL watchdog: a call did not return\n";
    (void)!write(1, msg, sizeof msg - 1);
    _exit(3);
}

static int is_codestream(const uint8_t* d, size_t n) {
    return d && n > 4 && d[0] == 0xFF && d[1] == 0x4F && d[n - 2] == 0xFF && d[n - 1] == 0xD9;
}

/* ---- input formats ----------------------------------------------------------------------- */
typedef struct {
    const char* name;
    gail_pixel_format fmt;
    int bits, sx, sy; /* 0 bits: interleaved (3 x 16-bit samples per pixel) */
} format_info;

static const format_info kFormats[] = {
    {"rgb48", GAIL_PIX_RGB48, 0, 0, 0},          {"xyz12", GAIL_PIX_XYZ12, 0, 0, 0},
    {"xyz12_msb", GAIL_PIX_XYZ12_MSB, 0, 0, 0},  {"yuv420p", GAIL_PIX_YUV420P, 8, 1, 1},
    {"yuv420p10", GAIL_PIX_YUV420P10, 10, 1, 1}, {"yuv420p12", GAIL_PIX_YUV420P12, 12, 1, 1},
    {"yuv422p", GAIL_PIX_YUV422P, 8, 1, 0},      {"yuv422p10", GAIL_PIX_YUV422P10, 10, 1, 0},
    {"yuv422p12", GAIL_PIX_YUV422P12, 12, 1, 0}, {"yuv444p", GAIL_PIX_YUV444P, 8, 0, 0},
    {"yuv444p10", GAIL_PIX_YUV444P10, 10, 0, 0}, {"yuv444p12", GAIL_PIX_YUV444P12, 12, 0, 0},
    {"gbrp", GAIL_PIX_GBRP, 8, 0, 0},            {"gbrp10", GAIL_PIX_GBRP10, 10, 0, 0},
    {"gbrp12", GAIL_PIX_GBRP12, 12, 0, 0},       {"gbrp16", GAIL_PIX_GBRP16, 16, 0, 0},
    {"xyz12_planar", GAIL_PIX_XYZ12_PLANAR, 12, 0, 0},
};

static const format_info* find_format(const char* name) {
    for (size_t i = 0; i < sizeof kFormats / sizeof kFormats[0]; ++i)
        if (!strcmp(kFormats[i].name, name)) return &kFormats[i];
    return NULL;
}

/* A picture of one frame: smooth gradients, moving with f, plus a little hashed noise so
 * that every code-block has something to code. Planes back to back in `buf`. */
static uint32_t noise(uint32_t x, uint32_t y, uint32_t c, uint32_t f) {
    uint32_t h = x * 73856093u ^ y * 19349663u ^ c * 83492791u ^ f * 2654435761u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    return h ^ (h >> 15);
}

typedef struct {
    uint8_t* buf;
    size_t bytes;
    const void* data[3];
} picture;

static void make_picture(picture* p, const format_info* fi, int w, int h, int f) {
    if (fi->bits == 0) {
        p->bytes = (size_t)w * h * 6;
        p->buf = (uint8_t*)malloc(p->bytes);
        uint16_t* px = (uint16_t*)p->buf;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c) {
                    const uint32_t v = ((uint32_t)(x + 7 * f) * (c + 1) + (uint32_t)y * 3 + noise(x, y, c, f) % 97) % 4096u;
                    px[((size_t)y * w + x) * 3 + c] =
                        (uint16_t)(fi->fmt == GAIL_PIX_XYZ12 ? v : fi->fmt == GAIL_PIX_XYZ12_MSB ? v << 4 : v * 16u + (v >> 8));
                }
        p->data[0] = p->buf;
        p->data[1] = p->data[2] = NULL;
        return;
    }
    const size_t bps = fi->bits > 8 ? 2 : 1;
    const int cw = (w + fi->sx) >> fi->sx, ch = (h + fi->sy) >> fi->sy;
    p->bytes = ((size_t)w * h + 2 * (size_t)cw * ch) * bps;
    p->buf = (uint8_t*)malloc(p->bytes);
    uint8_t* q = p->buf;
    const uint32_t maxv = (1u << fi->bits) - 1;
    for (int c = 0; c < 3; ++c) {
        const int pw = c ? cw : w, ph = c ? ch : h;
        const int full = fi->fmt >= GAIL_PIX_GBRP || !c; /* G'B'R' planes and luma span the range */
        for (int y = 0; y < ph; ++y)
            for (int x = 0; x < pw; ++x) {
                uint32_t v;
                if (full)
                    v = ((uint32_t)(x * 3 + y + f * 5 + c * 200) * (maxv / 255) + noise(x, y, c, f) % (maxv / 64 + 1)) % (maxv + 1);
                else
                    v = maxv / 2 + ((uint32_t)(x + y + f) % 64) * (maxv / 255) + noise(x, y, c, f) % (maxv / 128 + 1);
                if (fi->fmt == GAIL_PIX_XYZ12_PLANAR) v &= 4095u;
                if (bps == 1) q[(size_t)y * pw + x] = (uint8_t)v;
                else ((uint16_t*)q)[(size_t)y * pw + x] = (uint16_t)v;
            }
        p->data[c] = q;
        q += (size_t)pw * ph * bps;
    }
}

static void write_file(const char* dir, const char* what, int f, const void* data, size_t n) {
    if (!dir) return;
    char path[1024];
    snprintf(path, sizeof path, "%s/%s_%03d.bin", dir, what, f);
    FILE* fp = fopen(path, "wb");
    CHECK(fp != NULL, "cannot write %s", path);
    if (!fp) return;
    CHECK(fwrite(data, 1, n, fp) == n, "short write to %s", path);
    fclose(fp);
}

static void base_config(gail_config* cfg, const format_info* fi) {
    gail_config_defaults(cfg);
    cfg->profile = GAIL_PROFILE_2K;
    cfg->container = GAIL_CONTAINER_FLAT;
    cfg->fps = 24.0;
    cfg->max_mbps = 250.0;
    cfg->input.format = fi->fmt;
    cfg->frames_in_flight = 2;
    cfg->check = 1;
}

static gail_encoder* make(const char* name, gail_config* cfg) {
    gail_encoder* e = NULL;
    const int r = gail_encoder_create(cfg, &e);
    CHECK(r == GAIL_OK && e, "%s: create: %s (%s)", name, gail_status_string(r), gail_last_error());
    return r == GAIL_OK ? e : NULL;
}

static void picture_size(gail_encoder* e, const gail_config* cfg, int* w, int* h) {
    *w = cfg->input.width;
    *h = cfg->input.height;
    if (!*w || !*h) gail_encoder_frame_size(e, w, h);
}

/* ---- pull ------------------------------------------------------------------------------- */
/* pull <format> <frames> <outdir|-> [picture_w picture_h [frames_in_flight [chroma [full_range [matrix]]]]] */
static int test_pull(int argc, char** argv) {
    if (argc < 3) return 2;
    const format_info* fi = find_format(argv[0]);
    if (!fi) return 2;
    const int frames = atoi(argv[1]);
    const char* dir = strcmp(argv[2], "-") ? argv[2] : NULL;
    gail_config cfg;
    base_config(&cfg, fi);
    if (argc >= 5) {
        cfg.input.width = atoi(argv[3]);
        cfg.input.height = atoi(argv[4]);
    }
    if (argc >= 6) cfg.frames_in_flight = atoi(argv[5]);
    if (argc >= 7) cfg.input.chroma_location = (gail_chroma_location)atoi(argv[6]);
    if (argc >= 8) cfg.input.full_range = atoi(argv[7]);
    if (argc >= 9) cfg.input.matrix = (gail_matrix)atoi(argv[8]);
    gail_encoder* e = make("pull", &cfg);
    if (!e) return 1;
    int w, h;
    picture_size(e, &cfg, &w, &h);
    int64_t got = 0;
    size_t bytes = 0;
    for (int f = 0; f <= frames; ++f) {
        if (f < frames) {
            picture p;
            make_picture(&p, fi, w, h, f);
            write_file(dir, "in", f, p.buf, p.bytes);
            gail_frame fr;
            memset(&fr, 0, sizeof fr);
            for (int c = 0; c < 3; ++c) fr.data[c] = p.data[c];
            fr.opaque = (void*)(intptr_t)(1000 + f);
            CHECK(gail_encoder_send(e, &fr) == GAIL_OK, "pull: send %d: %s", f, gail_last_error());
            free(p.buf); /* the library copied the frame */
        } else {
            CHECK(gail_encoder_send(e, NULL) == GAIL_OK, "pull: end of stream");
        }
        for (;;) { /* whatever is ready; blocking at the end */
            gail_packet pkt;
            const int r = gail_encoder_receive(e, &pkt, f == frames);
            if (r == GAIL_EAGAIN || r == GAIL_EOF) break;
            CHECK(r == GAIL_OK, "pull: receive: %s", gail_last_error());
            if (r != GAIL_OK) break;
            CHECK(pkt.index == got && pkt.opaque == (void*)(intptr_t)(1000 + got), "pull: packet %lld out of order",
                  (long long)pkt.index);
            CHECK(pkt.compliant && is_codestream(pkt.data, pkt.size), "pull: packet %lld not a compliant codestream",
                  (long long)pkt.index);
            write_file(dir, "cs", (int)pkt.index, pkt.data, pkt.size);
            bytes += pkt.size;
            ++got;
            gail_packet_free(&pkt);
        }
    }
    gail_packet pkt;
    CHECK(gail_encoder_receive(e, &pkt, 1) == GAIL_EOF, "pull: no EOF");
    CHECK(got == frames, "pull: %lld of %d packets", (long long)got, frames);
    gail_frame fr;
    memset(&fr, 0, sizeof fr);
    fr.data[0] = fr.data[1] = fr.data[2] = &cfg;
    CHECK(gail_encoder_send(e, &fr) == GAIL_ERROR_STATE, "pull: send after the end accepted");
    CHECK(gail_encoder_encode(e, &fr, &pkt) == GAIL_ERROR_STATE, "pull: encode after the end accepted");
    CHECK(gail_encoder_finish(e) == GAIL_OK, "pull: finish after the end: %s", gail_last_error());
    printf("  pull %-12s %dx%d on %s: %lld packets, %zu bytes\n", fi->name, w, h, gail_encoder_devices(e),
           (long long)got, bytes);
    gail_encoder_destroy(e);
    return 0;
}

/* ---- callback --------------------------------------------------------------------------- */
typedef struct {
    int64_t next;
    int bad;
    size_t bytes;
} sink;

static void on_packet(void* user, const gail_packet* p) {
    sink* s = (sink*)user;
    if (p->index != s->next || p->opaque != (void*)(intptr_t)(500 + p->index) || !p->compliant ||
        !is_codestream(p->data, p->size))
        s->bad++;
    s->bytes += p->size;
    s->next++;
}

static int test_callback(int frames) {
    const format_info* fi = find_format("xyz12");
    gail_config cfg;
    base_config(&cfg, fi);
    sink s = {0, 0, 0};
    cfg.on_packet = on_packet;
    cfg.user = &s;
    gail_encoder* e = make("callback", &cfg);
    if (!e) return 1;
    int w, h;
    picture_size(e, &cfg, &w, &h);
    for (int f = 0; f < frames; ++f) {
        picture p;
        make_picture(&p, fi, w, h, f);
        gail_frame fr;
        memset(&fr, 0, sizeof fr);
        fr.data[0] = p.buf;
        fr.opaque = (void*)(intptr_t)(500 + f);
        CHECK(gail_encoder_send(e, &fr) == GAIL_OK, "callback: send: %s", gail_last_error());
        free(p.buf);
    }
    CHECK(gail_encoder_finish(e) == GAIL_OK, "callback: finish: %s", gail_last_error());
    CHECK(s.next == frames && s.bad == 0, "callback: %lld callbacks, %d bad", (long long)s.next, s.bad);
    gail_packet pkt;
    CHECK(gail_encoder_receive(e, &pkt, 0) == GAIL_ERROR_STATE, "callback: receive allowed in callback mode");
    printf("  callback: %d callbacks in order, %zu bytes\n", (int)s.next, s.bytes);
    gail_encoder_destroy(e);
    return 0;
}

/* ---- blocking encodes from several threads ------------------------------------------------ */
#define THREADS 4
#define PER_THREAD 3
typedef struct {
    gail_encoder* e;
    int id, w, h, ok;
    uint64_t sums[PER_THREAD];
} job;

static uint64_t fnv(const uint8_t* d, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) h = (h ^ d[i]) * 1099511628211ull;
    return h;
}

static void* worker(void* arg) {
    job* j = (job*)arg;
    const format_info* fi = find_format("xyz12");
    for (int k = 0; k < PER_THREAD; ++k) {
        picture p;
        make_picture(&p, fi, j->w, j->h, (j->id % 2) * 10 + k);
        gail_frame fr;
        memset(&fr, 0, sizeof fr);
        fr.data[0] = p.buf;
        fr.opaque = (void*)(intptr_t)(j->id * 100 + k);
        gail_packet pkt;
        memset(&pkt, 0, sizeof pkt);
        const int r = gail_encoder_encode(j->e, &fr, &pkt);
        if (r == GAIL_OK && pkt.opaque == fr.opaque && pkt.compliant && is_codestream(pkt.data, pkt.size)) {
            j->ok++;
            j->sums[k] = fnv(pkt.data, pkt.size);
        }
        gail_packet_free(&pkt);
        free(p.buf);
    }
    return NULL;
}

static int test_threads(int frames_in_flight) {
    const format_info* fi = find_format("xyz12");
    gail_config cfg;
    base_config(&cfg, fi);
    cfg.frames_in_flight = frames_in_flight;
    gail_encoder* e = make("threads", &cfg);
    if (!e) return 1;
    job jobs[THREADS];
    pthread_t th[THREADS];
    for (int i = 0; i < THREADS; ++i) {
        memset(&jobs[i], 0, sizeof jobs[i]);
        jobs[i].e = e;
        jobs[i].id = i;
        picture_size(e, &cfg, &jobs[i].w, &jobs[i].h);
        pthread_create(&th[i], NULL, worker, &jobs[i]);
    }
    int ok = 0;
    for (int i = 0; i < THREADS; ++i) {
        pthread_join(th[i], NULL);
        ok += jobs[i].ok;
    }
    CHECK(ok == THREADS * PER_THREAD, "threads: %d of %d frames came back right", ok, THREADS * PER_THREAD);
    /* threads 0 and 1 coded different pictures */
    int differ = 0;
    for (int k = 0; k < PER_THREAD; ++k)
        if (jobs[0].sums[k] == jobs[1].sums[k]) ++differ;
    CHECK(differ == 0, "threads: different pictures gave the same codestream");
    printf("  threads: %d threads x %d frames, each got its own codestream: %d ok\n", THREADS, PER_THREAD, ok);
    gail_encoder_destroy(e);
    return 0;
}

/* ---- cancel ----------------------------------------------------------------------------- */
typedef struct {
    gail_encoder* e;
    int w, h, sent, last, mode; /* mode 0: send forever; 1: encode forever */
} feeder;

static void* feed(void* arg) {
    feeder* f = (feeder*)arg;
    picture p;
    make_picture(&p, find_format("xyz12"), f->w, f->h, 0);
    for (;;) {
        gail_frame fr;
        memset(&fr, 0, sizeof fr);
        fr.data[0] = p.buf;
        gail_packet pkt;
        f->last = f->mode ? gail_encoder_encode(f->e, &fr, &pkt) : gail_encoder_send(f->e, &fr);
        if (f->last != GAIL_OK) break;
        if (f->mode) gail_packet_free(&pkt);
        f->sent++;
    }
    free(p.buf);
    return NULL;
}

typedef struct {
    gail_encoder* e;
    int calls;
} canceler;

static void cancel_in_callback(void* user, const gail_packet* p) {
    canceler* c = (canceler*)user;
    if (c->calls++ == 1) gail_encoder_cancel(c->e);
    (void)p;
}

static int test_cancel(void) {
    const format_info* fi = find_format("xyz12");
    for (int mode = 0; mode < 2; ++mode) {
        const char* name = mode ? "cancel encode" : "cancel send";
        gail_config cfg;
        base_config(&cfg, fi);
        gail_encoder* e = make(name, &cfg);
        if (!e) return 1;
        feeder f[2];
        pthread_t th[2];
        for (int i = 0; i < 2; ++i) {
            memset(&f[i], 0, sizeof f[i]);
            f[i].e = e;
            f[i].mode = mode;
            picture_size(e, &cfg, &f[i].w, &f[i].h);
            pthread_create(&th[i], NULL, feed, &f[i]);
        }
        sleep_ms(300);
        const double t0 = now_ms();
        gail_encoder_cancel(e);
        for (int i = 0; i < 2; ++i) pthread_join(th[i], NULL);
        const double woken = now_ms() - t0;
        for (int i = 0; i < 2; ++i)
            CHECK(f[i].last == GAIL_ERROR_CANCELLED, "%s: feeder %d ended with %s", name, i, gail_status_string(f[i].last));
        gail_packet pkt;
        gail_frame fr;
        memset(&fr, 0, sizeof fr);
        fr.data[0] = &cfg;
        CHECK(gail_encoder_receive(e, &pkt, 1) == GAIL_ERROR_CANCELLED, "%s: receive after cancel", name);
        CHECK(gail_encoder_send(e, &fr) == GAIL_ERROR_CANCELLED, "%s: send after cancel", name);
        CHECK(gail_encoder_encode(e, &fr, &pkt) == GAIL_ERROR_CANCELLED, "%s: encode after cancel", name);
        CHECK(gail_encoder_finish(e) == GAIL_ERROR_CANCELLED, "%s: finish after cancel", name);
        CHECK(gail_encoder_error(e) == NULL, "%s: cancel reported as an error", name);
        gail_encoder_cancel(e); /* twice is fine */
        gail_encoder_destroy(e);
        const double total = now_ms() - t0;
        printf("  %-14s %d+%d frames in, callers woken in %.0f ms, destroyed after %.0f ms\n", name, f[0].sent, f[1].sent,
               woken, total);
        CHECK(woken < 2000, "%s: callers took %.0f ms to wake", name, woken);
    }

    /* from inside on_packet */
    gail_config cfg;
    base_config(&cfg, fi);
    canceler c = {NULL, 0};
    cfg.on_packet = cancel_in_callback;
    cfg.user = &c;
    const int r = gail_encoder_create(&cfg, &c.e);
    CHECK(r == GAIL_OK, "cancel cb: create: %s", gail_last_error());
    if (r != GAIL_OK) return 1;
    int w, h;
    picture_size(c.e, &cfg, &w, &h);
    picture p;
    make_picture(&p, fi, w, h, 0);
    int rs = GAIL_OK, sent = 0;
    for (; sent < 40 && rs == GAIL_OK; ++sent) {
        gail_frame fr;
        memset(&fr, 0, sizeof fr);
        fr.data[0] = p.buf;
        rs = gail_encoder_send(c.e, &fr);
    }
    CHECK(gail_encoder_finish(c.e) == GAIL_ERROR_CANCELLED, "cancel cb: finish: %s", gail_last_error());
    CHECK(c.calls >= 2 && c.calls < 40, "cancel cb: %d callbacks", c.calls);
    printf("  %-14s canceled in the 2nd callback: %d callbacks, %d sends\n", "cancel cb", c.calls, sent);
    free(p.buf);
    gail_encoder_destroy(c.e);
    return 0;
}

/* ---- a failing device (stand-in plugin: GAIL_STUB_FAIL_AT) -------------------------------- */
/* lost <fail_at>: frames are encoded from two threads and sent from a third; once the device
   fails every caller returns FAILED (none hangs) and the error is reported */
static int test_lost(void) {
    const format_info* fi = find_format("xyz12");
    gail_config cfg;
    base_config(&cfg, fi);
    gail_encoder* e = make("lost", &cfg);
    if (!e) return 1;
    feeder f[3];
    pthread_t th[3];
    for (int i = 0; i < 3; ++i) {
        memset(&f[i], 0, sizeof f[i]);
        f[i].e = e;
        f[i].mode = i < 2;
        picture_size(e, &cfg, &f[i].w, &f[i].h);
        pthread_create(&th[i], NULL, feed, &f[i]);
    }
    for (int i = 0; i < 3; ++i) pthread_join(th[i], NULL);
    for (int i = 0; i < 3; ++i)
        CHECK(f[i].last == GAIL_ERROR_FAILED, "lost: feeder %d ended with %s", i, gail_status_string(f[i].last));
    gail_packet pkt;
    int r;
    while ((r = gail_encoder_receive(e, &pkt, 1)) == GAIL_OK) gail_packet_free(&pkt);
    CHECK(r == GAIL_ERROR_FAILED, "lost: receive ended with %s", gail_status_string(r));
    CHECK(gail_encoder_finish(e) == GAIL_ERROR_FAILED, "lost: finish");
    const char* why = gail_encoder_error(e);
    CHECK(why && strstr(why, "device lost"), "lost: error text %s", why ? why : "(none)");
    printf("  lost: %d+%d encoded, %d sent before the device failed: %s\n", f[0].sent, f[1].sent, f[2].sent,
           why ? why : "");
    gail_encoder_destroy(e);
    return 0;
}

/* ---- creation failures -------------------------------------------------------------------- */
/* create <expected status> [substring of the message] [frames_in_flight]: what creating a
   2K encoder returns (the loader's refusals, no device, no device memory) */
static int test_create(int argc, char** argv) {
    if (argc < 1) return 2;
    const int want = atoi(argv[0]);
    gail_config cfg;
    base_config(&cfg, find_format("xyz12"));
    if (argc >= 3) cfg.frames_in_flight = atoi(argv[2]);
    gail_encoder* e = NULL;
    const int r = gail_encoder_create(&cfg, &e);
    const char* msg = gail_last_error();
    CHECK(r == want, "create returned %s (%d), expected %d: %s", gail_status_string(r), r, want, msg);
    if (want != GAIL_OK) CHECK(e == NULL, "create failed but gave an encoder");
    if (argc >= 2 && want != GAIL_OK) CHECK(strstr(msg, argv[1]) != NULL, "message \"%s\" lacks \"%s\"", msg, argv[1]);
    gail_device_info dev[4];
    const int n = gail_enumerate_devices(GAIL_BACKEND_AUTO, dev, 4);
    printf("  create: %s (%s); %d device(s)%s%s\n", gail_status_string(r), r == GAIL_OK ? "" : msg, n, n > 0 ? ": " : "",
           n > 0 ? dev[0].name : "");
    if (e) {
        /* the encoder made with fewer slots still codes */
        picture p;
        int w, h;
        picture_size(e, &cfg, &w, &h);
        make_picture(&p, find_format("xyz12"), w, h, 3);
        gail_frame fr;
        memset(&fr, 0, sizeof fr);
        fr.data[0] = p.buf;
        gail_packet pkt;
        CHECK(gail_encoder_encode(e, &fr, &pkt) == GAIL_OK, "encode: %s", gail_last_error());
        gail_packet_free(&pkt);
        free(p.buf);
        gail_encoder_destroy(e);
    }
    return 0;
}

/* ---- served form ------------------------------------------------------------------------ */
#include <dlfcn.h>

typedef int (*stub_control_fn)(const char*, long);

/* The stand-in plugin's knob setter, when the plugin next to libgail.so is the stand-in. */
static stub_control_fn stub_control(void) {
    Dl_info info;
    if (!dladdr((void*)&gail_version, &info) || !info.dli_fname) return NULL;
    char path[1024];
    snprintf(path, sizeof path, "%s", info.dli_fname);
    char* slash = strrchr(path, '/');
    if (!slash) return NULL;
    snprintf(slash + 1, sizeof path - (size_t)(slash + 1 - path), "libgail_hip.so");
    void* h = dlopen(path, RTLD_NOW | RTLD_NOLOAD);
    return h ? (stub_control_fn)dlsym(h, "gail_stub_control") : NULL;
}

typedef struct {
    pthread_mutex_t m;
    int done, ok, nomem, failed, compliant;
    int order[256];
} server_sink;

static void on_result(void* user, const gail_server_result* r) {
    server_sink* s = (server_sink*)user;
    pthread_mutex_lock(&s->m);
    if (s->done < 256) s->order[s->done] = (int)r->id;
    s->done++;
    if (r->status == GAIL_OK) {
        s->ok++;
        if (r->packet && r->packet->compliant && is_codestream(r->packet->data, r->packet->size)) s->compliant++;
    } else if (r->status == GAIL_ERROR_NOMEM) {
        s->nomem++;
    } else {
        s->failed++;
    }
    pthread_mutex_unlock(&s->m);
}

static void stream_config(gail_config* c, int fourk) {
    gail_config_defaults(c);
    c->profile = fourk ? GAIL_PROFILE_4K : GAIL_PROFILE_2K;
    c->container = GAIL_CONTAINER_FLAT;
    c->fps = 24.0;
    c->max_mbps = 250.0;
    c->input.format = GAIL_PIX_XYZ12_PLANAR;
}

/* One X'Y'Z' planar frame per size, shared by every submit (the server reads it before the
   frame's result arrives; nothing writes it). */
static uint16_t* g_frames[2];

static int submit(gail_server* srv, uint64_t id, int fourk) {
    gail_config c;
    stream_config(&c, fourk);
    int w, h;
    gail_container_size(c.profile, c.container, &w, &h);
    if (!g_frames[fourk]) {
        g_frames[fourk] = (uint16_t*)malloc((size_t)w * h * 6);
        for (size_t i = 0; i < (size_t)w * h * 3; ++i) g_frames[fourk][i] = (uint16_t)(noise((uint32_t)i, 7, 1, fourk) & 4095u);
    }
    gail_frame fr;
    memset(&fr, 0, sizeof fr);
    for (int p = 0; p < 3; ++p) fr.data[p] = g_frames[fourk] + (size_t)p * w * h;
    return gail_server_submit(srv, id, fourk ? 2 : 1, &c, &fr, 0.0);
}

static gail_server* make_server(server_sink* sink, int slots, int jobs, uint64_t budget) {
    gail_server_config sc;
    gail_server_config_init(&sc, sizeof sc);
    sc.slots_per_job = slots;
    sc.max_jobs = jobs;
    sc.memory_budget = budget;
    sc.check = 1;
    sc.cpu_threads = 3;
    sc.host_rate_control = 1;
    sc.on_result = on_result;
    sc.user = sink;
    memset(sink, 0, sizeof *sink);
    pthread_mutex_init(&sink->m, NULL);
    gail_server* srv = NULL;
    const int r = gail_server_create(&sc, &srv);
    CHECK(r == GAIL_OK && srv, "server create: %s (%s)", gail_status_string(r), gail_last_error());
    return srv;
}

static void print_stats(const char* name, gail_server* srv) {
    gail_server_stats st;
    memset(&st, 0, sizeof st);
    gail_server_get_stats(srv, &st);
    printf("  %-16s %llu frames (%llu failed); jobs now %d (2K %d, 4K %d), made %d, retired %d, most %d; "
           "in flight at most %d; peak %.0f MiB\n",
           name, (unsigned long long)st.frames, (unsigned long long)st.failed, st.jobs, st.jobs_2k, st.jobs_4k,
           st.jobs_made, st.jobs_retired, st.max_jobs, st.max_frames_in_flight, st.peak_job_bytes / 1048576.0);
}

/* server <case>: mixed | classes | pressure | nomem | lost (stand-in plugin only, but for mixed) */
static int test_server(const char* which) {
    const uint64_t MB = 1ull << 20;
    server_sink sink;
    gail_device_info info[1];
    gail_enumerate_devices(GAIL_BACKEND_AUTO, info, 1);  /* loads the plugin */
    stub_control_fn knob = stub_control();
    if (!strcmp(which, "mixed")) {
        /* frames of both sizes arriving together, with room for a job per two frames: each
           class gets jobs of its own slot size (none has to be retired for the other) */
        gail_server* srv = make_server(&sink, 2, 8, 0);
        if (!srv) return 1;
        for (int i = 0; i < 16; ++i) CHECK(submit(srv, i, i % 2) == GAIL_OK, "submit %d: %s", i, gail_last_error());
        gail_server_drain(srv);
        gail_server_stats st;
        gail_server_get_stats(srv, &st);
        print_stats("mixed", srv);
        CHECK(sink.done == 16 && sink.ok == 16 && sink.compliant == 16, "mixed: %d results, %d ok, %d compliant",
              sink.done, sink.ok, sink.compliant);
        CHECK(st.jobs_2k >= 1 && st.jobs_4k >= 1, "mixed: no job of each slot class");
        CHECK(st.slot_bytes_2k > 0 && st.slot_bytes_2k < st.slot_bytes, "mixed: 2K slots not smaller than 4K slots");
        gail_server_destroy(srv);
        return 0;
    }
    if (!knob) {
        printf("  server %s needs the stand-in plugin\n", which);
        return 2;
    }
    if (!strcmp(which, "classes")) {
        /* a budget of two 2K jobs: a 4K frame retires the idle 2K jobs to make its own */
        gail_server* probe = make_server(&sink, 2, 4, 0);
        if (!probe) return 1;
        gail_server_stats st;
        gail_server_get_stats(probe, &st);
        gail_server_destroy(probe);
        const uint64_t job2k = 2 * st.slot_bytes_2k, job4k = 2 * st.slot_bytes;
        gail_server* srv = make_server(&sink, 2, 4, job4k + job2k / 2);
        if (!srv) return 1;
        for (int i = 0; i < 6; ++i) CHECK(submit(srv, i, 0) == GAIL_OK, "submit 2K %d", i);
        gail_server_drain(srv);
        gail_server_get_stats(srv, &st);
        CHECK(st.jobs_2k >= 2, "classes: 2K frames made %d 2K jobs", st.jobs_2k);
        for (int i = 0; i < 2; ++i) CHECK(submit(srv, 100 + i, 1) == GAIL_OK, "submit 4K %d", i);
        gail_server_drain(srv);
        gail_server_get_stats(srv, &st);
        print_stats("classes", srv);
        CHECK(sink.ok == 8 && sink.compliant == 8, "classes: %d ok of 8", sink.ok);
        CHECK(st.jobs_retired >= 1 && st.jobs_4k == 1, "classes: %d retired, %d 4K jobs", st.jobs_retired, st.jobs_4k);
        CHECK(st.peak_job_bytes <= job4k + job2k / 2, "classes: %llu bytes held over the budget",
              (unsigned long long)st.peak_job_bytes);
        gail_server_destroy(srv);
        return 0;
    }
    if (!strcmp(which, "pressure")) {
        /* another process takes the device's memory: idle jobs are given up one at a time,
           the last one kept, and frames still code */
        knob("GAIL_STUB_MEM_MB", 16 * 1024);
        gail_server* srv = make_server(&sink, 2, 4, 0);
        if (!srv) return 1;
        for (int i = 0; i < 16; ++i) CHECK(submit(srv, i, 1) == GAIL_OK, "submit %d", i);
        gail_server_drain(srv);
        gail_server_stats st;
        gail_server_get_stats(srv, &st);
        const int before = st.jobs;
        CHECK(before >= 2, "pressure: only %d jobs made", before);
        knob("GAIL_STUB_MEM_USED_MB", (long)(16 * 1024 - st.job_bytes / MB - 100));
        for (int i = 0; i < 12; ++i) {
            CHECK(submit(srv, 100 + i, 1) == GAIL_OK, "submit under pressure %d", i);
            gail_server_drain(srv);
            sleep_ms(300);
        }
        gail_server_get_stats(srv, &st);
        print_stats("pressure", srv);
        CHECK(st.jobs_retired >= 1 && st.jobs >= 1 && st.jobs < before, "pressure: %d of %d jobs left, %d retired",
              st.jobs, before, st.jobs_retired);
        CHECK(sink.ok == 28, "pressure: %d of 28 frames ok", sink.ok);
        gail_server_destroy(srv);
        knob("GAIL_STUB_MEM_USED_MB", 0);
        return 0;
    }
    if (!strcmp(which, "nomem")) {
        /* a budget that holds 2K jobs but no 4K job: 4K frames fail with NOMEM, 2K frames code */
        gail_server* probe = make_server(&sink, 1, 2, 0);
        if (!probe) return 1;
        gail_server_stats st;
        gail_server_get_stats(probe, &st);
        gail_server_destroy(probe);
        gail_server* srv = make_server(&sink, 1, 2, st.slot_bytes_2k * 3 / 2);
        if (!srv) return 1;
        for (int i = 0; i < 4; ++i) CHECK(submit(srv, i, i % 2) == GAIL_OK, "submit %d", i);
        gail_server_drain(srv);
        print_stats("nomem", srv);
        CHECK(sink.ok == 2 && sink.nomem == 2, "nomem: %d ok, %d out of memory", sink.ok, sink.nomem);
        gail_server_destroy(srv);
        return 0;
    }
    if (!strcmp(which, "lost")) {
        /* the device fails mid-stream: every frame still gets a result, and destroy returns */
        knob("GAIL_STUB_FAIL_AT", 5);
        gail_server* srv = make_server(&sink, 2, 2, 0);
        if (!srv) return 1;
        for (int i = 0; i < 12; ++i) CHECK(submit(srv, i, i % 2) == GAIL_OK, "submit %d", i);
        gail_server_drain(srv);
        print_stats("lost", srv);
        CHECK(sink.done == 12 && sink.failed >= 1 && sink.ok >= 1, "lost: %d results, %d ok, %d failed", sink.done,
              sink.ok, sink.failed);
        gail_server_destroy(srv);
        knob("GAIL_STUB_FAIL_AT", 0);
        return 0;
    }
    return 2;
}

/* ---- configuration checks (binary compatibility rules of gail.h) -------------------------- */
static int test_config(void) {
    gail_encoder* e = NULL;
    struct {
        gail_config c;
        uint64_t later[4];
    } newer;
    memset(&newer, 0xAB, sizeof newer);
    gail_config_init(&newer.c, sizeof newer);
    CHECK(newer.c.struct_size == sizeof newer && newer.later[0] == 0 && newer.later[3] == 0 && newer.c.fps == 24,
          "gail_config_init with a larger size");
    newer.c.input.format = GAIL_PIX_XYZ12;
    CHECK(gail_encoder_create(&newer.c, &e) == GAIL_OK, "newer header, unused fields: %s", gail_last_error());
    gail_encoder_destroy(e);
    e = NULL;
    newer.later[1] = 7;
    CHECK(gail_encoder_create(&newer.c, &e) == GAIL_ERROR_UNSUPPORTED && !e, "newer header, a setting we lack accepted");

    struct {
        gail_config c;
        uint64_t guard;
    } g;
    g.guard = 0x1234;
    gail_config_init(&g.c, sizeof g.c);
    CHECK(g.guard == 0x1234, "gail_config_init wrote past the struct");

    gail_config cfg;
    gail_config_defaults(&cfg);
    cfg.struct_size = offsetof(gail_config, log);
    CHECK(gail_encoder_create(&cfg, &e) == GAIL_ERROR_INVALID, "config smaller than 1.0 accepted");

    gail_config_defaults(&cfg);
    cfg.input.format = GAIL_PIX_XYZ12;
    cfg.input.reserved[0] = 1;
    CHECK(gail_encoder_create(&cfg, &e) == GAIL_ERROR_INVALID && strstr(gail_last_error(), "reserved"),
          "non-zero reserved field accepted");
    gail_config_defaults(&cfg);
    cfg.input.width = 5000;
    cfg.input.height = 100;
    CHECK(gail_encoder_create(&cfg, &e) == GAIL_ERROR_INVALID, "picture larger than the container accepted");
    gail_config_defaults(&cfg);
    cfg.backend = GAIL_BACKEND_CUDA;
    CHECK(gail_encoder_create(&cfg, &e) == GAIL_ERROR_UNSUPPORTED, "CUDA backend accepted");
    printf("  config: size and reserved-field rules hold\n");
    return 0;
}

int main(int argc, char** argv) {
    setvbuf(stdout, NULL, _IOLBF, 0);
    if (argc < 2) {
        fprintf(stderr, "usage: test_api pull|callback|threads|cancel|lost|create|config|server [arguments]\n");
        return 2;
    }
    signal(SIGALRM, on_alarm);
    alarm((unsigned)(getenv("GAIL_TEST_WATCHDOG_S") ? atoi(getenv("GAIL_TEST_WATCHDOG_S")) : 240));
    const char* c = argv[1];
    int r;
    if (!strcmp(c, "pull")) r = test_pull(argc - 2, argv + 2);
    else if (!strcmp(c, "callback")) r = test_callback(argc > 2 ? atoi(argv[2]) : 6);
    else if (!strcmp(c, "threads")) r = test_threads(argc > 2 ? atoi(argv[2]) : 2);
    else if (!strcmp(c, "cancel")) r = test_cancel();
    else if (!strcmp(c, "lost")) r = test_lost();
    else if (!strcmp(c, "create")) r = test_create(argc - 2, argv + 2);
    else if (!strcmp(c, "config")) r = test_config();
    else if (!strcmp(c, "server") && argc > 2) r = test_server(argv[2]);
    else r = 2;
    if (r == 2) {
        fprintf(stderr, "test_api: bad arguments\n");
        return 2;
    }
    printf("%s: %d failure(s)\n", c, failures);
    return failures || r ? 1 : 0;
}
