'use strict';
// Test driver for the Node-API path: replays protocol messages (src/protocol.ts:29-45) read from
// <dir>/spec.json through createGpuWorker (meercadraytracer_amd/napi/gpu_worker.js) and writes
// every response: tile_<job>.bin (the Float32Array) and results.json (job -> 'tile' | 'error:...').
// Arrays in the spec are {"__array__": {"file", "type"}} references to raw files in <dir>.
const fs = require('fs');
const path = require('path');
const { createGpuWorker } = require(path.join(__dirname, '..', 'meercadraytracer_amd', 'napi', 'gpu_worker.js'));

const dir = process.argv[2];
const spec = JSON.parse(fs.readFileSync(path.join(dir, 'spec.json'), 'utf8'));

function array(ref) {
  const b = fs.readFileSync(path.join(dir, ref.file));
  const ab = b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength);
  return new globalThis[ref.type](ab);
}

function revive(v) {
  if (Array.isArray(v)) return v.map(revive);
  if (v && typeof v === 'object') {
    if (v.__array__) return array(v.__array__);
    const o = {};
    for (const k of Object.keys(v)) o[k] = revive(v[k]);
    return o;
  }
  return v;
}

const messages = revive(spec.messages);
const worker = createGpuWorker();
const results = {};
let pending = messages.filter(m => m.type === 'tile').length;
worker.addEventListener('message', event => {
  const r = event.data;
  if (r.type === 'tile') {
    fs.writeFileSync(path.join(dir, `tile_${r.job}.bin`), Buffer.from(r.data.buffer, r.data.byteOffset, r.data.byteLength));
    results[r.job] = 'tile';
  } else {
    results[r.job] = 'error:' + r.message;
  }
  if (--pending === 0) {
    fs.writeFileSync(path.join(dir, 'results.json'), JSON.stringify(results));
    worker.terminate();
  }
});
worker.addEventListener('error', event => {
  console.error('worker error: ' + event.message);
  process.exit(1);
});
for (const m of messages) worker.postMessage(m);
