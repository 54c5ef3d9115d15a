/*
 * CPU oracle — clip state machine and save/restore (test infrastructure only).
 *
 * Restates src/vkvg_context.c:698-795 (clip, reset_clip, clear) and :1251-1512 (save, restore)
 * on the oracle's 8-bit stencil: bit 0 FILL, bit 1 CLIP (a draw passes where CLIP == 0), bits
 * 2..7 the CLIP bits of up to six saved levels; every sixth level the whole stencil is first
 * copied into a spill image, which the matching restore copies back.
 */
#include "ovg_internal.h"

/* _get_previous_clip_state, src/vkvg_context.c:698-702 */
static int previous_clip_state(const octx *c) { return c->saved ? c->saved->clipState : OVG_CLIP_CLEAR; }

static void stencil_clear(octx *c) {
    memset(c->surf->stencil, 0, (size_t)c->surf->width * c->surf->height);
}

/* _update_cur_pattern for LINEAR/RADIAL, src/vkvg_context_internal.c:774-826: the control points
 * go through the inverse of the pattern's matrix (identity when it has none or it is singular,
 * :797-803), then through the CTM current at this call */
void o_set_gradient(octx *c, const ovg_grad *def) {
    ovg_grad g = *def;
    /* the target's size goes to the unused source field (:782-786) */
    c->srcRect[0] = (float)c->surf->width;
    c->srcRect[1] = (float)c->surf->height;
    c->srcRect[2] = c->srcRect[3] = 0.f;
    if (g.count < 2) {
        c->status = 10; /* VKVG_STATUS_PATTERN_INVALID_GRADIENT */
        return;
    }
    mat6 pm = g.matrix;
    if (g.hasMatrix) {
        if (m_invert(&pm) != 0)
            pm = (mat6){1, 0, 0, 1, 0, 0};
        m_transform_point(&pm, &g.cp[0][0], &g.cp[0][1]);
    }
    m_transform_point(&c->mat, &g.cp[0][0], &g.cp[0][1]);
    if (g.type == SRC_LINEAR) {
        if (g.hasMatrix)
            m_transform_point(&pm, &g.cp[0][2], &g.cp[0][3]);
        m_transform_point(&c->mat, &g.cp[0][2], &g.cp[0][3]);
    } else {
        if (g.hasMatrix) {
            m_transform_point(&pm, &g.cp[1][0], &g.cp[1][1]);
        }
        m_transform_point(&c->mat, &g.cp[1][0], &g.cp[1][1]);
        if (g.hasMatrix) { /* radii, :816-819 */
            m_transform_distance(&pm, &g.cp[0][2], &g.cp[0][3]);
            m_transform_distance(&pm, &g.cp[1][2], &g.cp[0][3]);
        }
        m_transform_distance(&c->mat, &g.cp[0][2], &g.cp[0][3]);
        m_transform_distance(&c->mat, &g.cp[1][2], &g.cp[0][3]); /* as written in the reference */
    }
    c->grad = g;
    c->srcType = g.type;
}

/* _update_cur_pattern for SURFACE, src/vkvg_context_internal.c:705-772: the source size, and the
 * pattern's matrix (not inverted) multiplied into the inverse matrix */
void o_set_surface_pattern(octx *c, const ovg_surf_pat *p) {
    c->surfPat = *p;
    c->srcType = SRC_SURFACE;
    c->srcRect[2] = (float)c->sources[p->index].width;
    c->srcRect[3] = (float)c->sources[p->index].height;
    if (p->hasMatrix)
        m_multiply(&c->matInv, &c->matInv, &p->matrix);
}

/* _set_mat_inv_and_vkCmdPush, src/vkvg_context_internal.c:672-676 */
void o_update_mat_inv(octx *c) {
    c->matInv = c->mat;
    m_invert(&c->matInv);
}

/* vkvg_clip_preserve, src/vkvg_context.c:971-976 -> _clip_preserve :754-795 */
void o_clip_preserve(octx *c) { r_clip(c); }

/* vkvg_reset_clip, src/vkvg_context.c:719-733; _reset_clip clears the whole stencil (:706-718) */
void o_reset_clip(octx *c) {
    if (c->clipState == OVG_CLIP_CLEAR)
        return;
    c->clipState = previous_clip_state(c) == OVG_CLIP_CLEAR ? OVG_CLIP_NONE : OVG_CLIP_CLEAR;
    stencil_clear(c);
}

/* vkvg_clear, src/vkvg_context.c:734-753: colour and stencil cleared */
void o_clear(octx *c) {
    c->clipState = previous_clip_state(c) == OVG_CLIP_CLEAR ? OVG_CLIP_NONE : OVG_CLIP_CLEAR;
    r_clear(c);
}

/* vkvg_save, src/vkvg_context.c:1251-1375 */
void o_save(octx *c) {
    ovg_surface *S = c->surf;
    const size_t np = (size_t)S->width * S->height;
    struct osave *sav = (struct osave *)calloc(1, sizeof(struct osave));
    if (c->clipState == OVG_CLIP_CLIP) {
        sav->clipState = OVG_CLIP_SAVED;
        if (c->curSavBit > 0 && c->curSavBit % 6 == 0) { /* a new spill image takes the stencil */
            uint32_t idx = c->curSavBit / 6 - 1;
            if (idx >= c->nSpills) {
                c->spills = (uint8_t **)realloc(c->spills, (idx + 1) * sizeof(uint8_t *));
                for (uint32_t i = c->nSpills; i <= idx; i++)
                    c->spills[i] = NULL;
                c->nSpills = idx + 1;
            }
            if (!c->spills[idx])
                c->spills[idx] = (uint8_t *)malloc(np);
            memcpy(c->spills[idx], S->stencil, np);
        }
        /* full-screen quad: reference CLIP|bit, compare mask CLIP, write mask bit -> bit := CLIP */
        const uint8_t bit = (uint8_t)(1u << (c->curSavBit % 6 + 2));
        for (size_t k = 0; k < np; k++)
            S->stencil[k] = (uint8_t)((S->stencil[k] & ~bit) | ((S->stencil[k] & 2) ? bit : 0));
        c->curSavBit++;
    } else if (c->clipState == OVG_CLIP_NONE) {
        sav->clipState = previous_clip_state(c) & 0x03;
    } else {
        sav->clipState = OVG_CLIP_CLEAR;
    }
    sav->dashOffset = c->dashOffset;
    sav->dashCount = c->dashCount;
    if (c->dashCount) {
        sav->dashes = (float *)malloc(c->dashCount * sizeof(float));
        memcpy(sav->dashes, c->dashes, c->dashCount * sizeof(float));
    }
    sav->lineWidth = c->lineWidth;
    sav->miterLimit = c->miterLimit;
    sav->op = c->op;
    sav->lineCap = c->lineCap;
    sav->fillRule = c->fillRule;
    sav->mat = c->mat; /* pushConsts: CTM, inverse, source rectangle, opacity */
    sav->matInv = c->matInv;
    memcpy(sav->srcRect, c->srcRect, sizeof(sav->srcRect));
    sav->opacity = c->opacity;
    sav->patId = c->patId;
    sav->patType = c->srcType;
    if (c->patId) {
        sav->gradDef = c->gradDef;
        sav->surfPat = c->surfPat;
    } else {
        sav->curColor = c->curColor;
    }
    sav->next = c->saved;
    c->saved = sav;
}

/* vkvg_restore, src/vkvg_context.c:1376-1512 */
void o_restore(octx *c) {
    ovg_surface *S = c->surf;
    const size_t np = (size_t)S->width * S->height;
    if (!c->saved) {
        c->status = 3; /* VKVG_STATUS_INVALID_RESTORE */
        return;
    }
    struct osave *sav = c->saved;
    c->saved = sav->next;
    c->mat = sav->mat;
    c->matInv = sav->matInv;
    memcpy(c->srcRect, sav->srcRect, sizeof(c->srcRect));
    c->opacity = sav->opacity;
    if (c->clipState != OVG_CLIP_NONE) {
        if (c->clipState == OVG_CLIP_CLIP && sav->clipState == OVG_CLIP_CLEAR) {
            stencil_clear(c);
        } else {
            /* full-screen quad: reference CLIP|bit, compare mask bit, write mask CLIP -> CLIP := bit.
             * (curSavBit - 1) % 6 + 2 in C int arithmetic: with curSavBit == 0 it is 1, the CLIP
             * bit itself, and the pass copies CLIP onto itself. */
            const int sh = ((int)c->curSavBit - 1) % 6 + 2;
            const uint8_t bit = (uint8_t)(1u << sh);
            for (size_t k = 0; k < np; k++)
                S->stencil[k] = (uint8_t)((S->stencil[k] & ~2u) | ((S->stencil[k] & bit) ? 2u : 0u));
        }
    }
    if (sav->clipState == OVG_CLIP_SAVED) {
        c->curSavBit--;
        if (c->curSavBit > 0 && c->curSavBit % 6 == 0) { /* the spill image comes back */
            uint32_t idx = c->curSavBit / 6 - 1;
            memcpy(S->stencil, c->spills[idx], np);
            free(c->spills[idx]);
            c->spills[idx] = NULL;
        }
    }
    c->clipState = OVG_CLIP_NONE;
    c->dashOffset = sav->dashOffset;
    free(c->dashes);
    c->dashes = NULL;
    c->dashCount = sav->dashCount;
    if (c->dashCount) {
        c->dashes = (float *)malloc(c->dashCount * sizeof(float));
        memcpy(c->dashes, sav->dashes, c->dashCount * sizeof(float));
    }
    c->lineWidth = sav->lineWidth;
    c->miterLimit = sav->miterLimit;
    c->op = sav->op;
    c->lineCap = sav->lineCap;
    c->lineJoin = 0; /* restored from a field vkvg_save never writes (calloc'd: MITER), :1492 */
    c->fillRule = sav->fillRule;
    if (sav->patId) {
        if (sav->patId != c->patId) { /* _update_cur_pattern with the restored push constants */
            c->patId = sav->patId;
            if (sav->patType == SRC_SURFACE) {
                o_set_surface_pattern(c, &sav->surfPat);
            } else {
                c->gradDef = sav->gradDef;
                o_set_gradient(c, &sav->gradDef);
            }
        }
    } else {
        c->curColor = sav->curColor;
        c->patId = 0;
        c->srcType = SRC_SOLID;
    }
    free(sav->dashes);
    free(sav);
}

void o_state_free(octx *c) {
    while (c->saved) {
        struct osave *s = c->saved;
        c->saved = s->next;
        free(s->dashes);
        free(s);
    }
    for (uint32_t i = 0; i < c->nSpills; i++)
        free(c->spills[i]);
    free(c->spills);
    c->spills = NULL;
    c->nSpills = 0;
}
