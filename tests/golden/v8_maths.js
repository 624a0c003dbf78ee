// Evaluates one of V8's fdlibm-port Math functions elementwise:
//   node v8_maths.js <log|exp|sin|cos> <in.f64> <out.f64>
// Both files hold little-endian IEEE doubles.
'use strict';
const fs = require('fs');
const [fn, inPath, outPath] = process.argv.slice(2);
const f = { log: Math.log, exp: Math.exp, sin: Math.sin, cos: Math.cos }[fn];
if (!f) throw new Error('unknown function ' + fn);
const buf = fs.readFileSync(inPath);
const x = new Float64Array(buf.buffer, buf.byteOffset, buf.length / 8);
const y = new Float64Array(x.length);
for (let i = 0; i < x.length; i++) y[i] = f(x[i]);
fs.writeFileSync(outPath, Buffer.from(y.buffer));
