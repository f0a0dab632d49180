// Camera messages of the session front end: the LookOut matrix a client
// sends (network space: meters, the data's own frame) and the modelview the
// renderer draws with (Livre space), converted as Livre's ZeroEQ communicator
// converts them (communicator.cpp:106-133 onCamera, 257-283 _createLookOut).
// Matrices are 16 floats, column-major (element (row r, column c) at
// [4 c + r]), as vmmlib stores Matrix4f and lexis stores LookOut. The
// transport (ZeroEQ publish / subscribe, HTTP) is not part of this library.
#include "lvr_internal.h"

#include <cmath>
#include <cstring>

namespace
{
struct M4
{
    float a[16];
    float& at(int r, int c) { return a[4 * c + r]; }
    float at(int r, int c) const { return a[4 * c + r]; }
};

struct V4
{
    float v[4];
};

M4 identity()
{
    M4 m;
    std::memset(m.a, 0, sizeof m.a);
    for (int i = 0; i < 4; ++i)
        m.at(i, i) = 1.f;
    return m;
}

V4 mul(const M4& m, const V4& x)
{
    V4 y;
    for (int r = 0; r < 4; ++r)
        y.v[r] = m.at(r, 0) * x.v[0] + m.at(r, 1) * x.v[1] + m.at(r, 2) * x.v[2] +
                 m.at(r, 3) * x.v[3];
    return y;
}

V4 neg(const V4& x)
{
    return V4{{-x.v[0], -x.v[1], -x.v[2], -x.v[3]}};
}

// Inverse by cofactors, in float; false for a singular matrix.
bool inverse(const M4& m, M4& out)
{
    const float* a = m.a;
    float inv[16];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] +
             a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] -
             a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] +
             a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] -
              a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] -
             a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] +
             a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] -
             a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] +
              a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] +
             a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] -
             a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] +
              a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] -
              a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] -
             a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] +
             a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] -
              a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] +
              a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    const float det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    if (det == 0.f || !std::isfinite(det))
        return false;
    for (int i = 0; i < 16; ++i)
        out.a[i] = inv[i] / det;
    return true;
}

// The 3x3 rotation of a view matrix in an otherwise identity 4x4 (a
// default-constructed vmmlib Matrix4f with its upper 3x3 set).
M4 rotationOf(const M4& m)
{
    M4 r = identity();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.at(i, j) = m.at(i, j);
    return r;
}

// A view matrix from a rotation and a translation column.
void compose(const M4& rotation, const V4& t, float out[16])
{
    M4 m = rotation;
    for (int i = 0; i < 3; ++i)
        m.at(i, 3) = t.v[i];
    m.at(3, 3) = 1.f;
    std::memcpy(out, m.a, sizeof m.a);
}

bool load(const float* in, M4& m)
{
    std::memcpy(m.a, in, sizeof m.a);
    for (float x : m.a)
        if (!std::isfinite(x))
            return false;
    return true;
}
} // namespace

extern "C" {

int lvr_lookout_to_modelview(const float lookout[16], const float data_to_livre[16],
                             float meter_to_data_unit, float out_modelview[16])
{
    M4 net, d = identity();
    if (!lookout || !out_modelview || !load(lookout, net) ||
        (data_to_livre && !load(data_to_livre, d)) || !std::isfinite(meter_to_data_unit) ||
        meter_to_data_unit == 0.f)
        return LVR_EINVAL;
    const M4 rotation = rotationOf(net);
    M4 rinv;
    if (!inverse(rotation, rinv))
        return LVR_EINVAL;
    V4 t{{net.at(0, 3), net.at(1, 3), net.at(2, 3), net.at(3, 3)}};
    // the camera position in network space, in data units
    t = neg(mul(rinv, t));
    for (float& x : t.v)
        x *= meter_to_data_unit;
    t.v[3] = 1.f;
    // into Livre space, then back into a view translation
    t = neg(mul(rotation, mul(d, t)));
    t.v[3] = 1.f;
    compose(rotation, t, out_modelview);
    return LVR_OK;
}

int lvr_modelview_to_lookout(const float modelview[16], const float data_to_livre[16],
                             float meter_to_data_unit, float out_lookout[16])
{
    M4 mv, d = identity();
    if (!modelview || !out_lookout || !load(modelview, mv) ||
        (data_to_livre && !load(data_to_livre, d)) || !std::isfinite(meter_to_data_unit) ||
        meter_to_data_unit == 0.f)
        return LVR_EINVAL;
    const M4 rotation = rotationOf(mv);
    M4 rinv, dinv;
    if (!inverse(rotation, rinv) || !inverse(d, dinv))
        return LVR_EINVAL;
    V4 t{{mv.at(0, 3), mv.at(1, 3), mv.at(2, 3), mv.at(3, 3)}};
    t = neg(mul(rinv, t));
    t.v[3] = 1.f;
    t = neg(mul(rotation, mul(dinv, t)));
    for (float& x : t.v)
        x *= 1.0f / meter_to_data_unit;
    t.v[3] = 1.f;
    compose(rotation, t, out_lookout);
    return LVR_OK;
}
}
